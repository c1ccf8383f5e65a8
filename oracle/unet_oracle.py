"""CPU oracle of the warm-start network (test infrastructure: only tests/ and
bench.py's checks import it).

The same four-level encoder-decoder as mantiis_amd/csrc/mts_net.hip: 3x3
periodic convolutions with ReLU, 2x2 average pooling, nearest 2x upsampling
concatenated with the skip tensor, a 1x1 head, and the masked mean-square
loss sum m |o - t|^2 / (2 sum m).  Parameters are one flat array laid out as
the library's (per layer: weights [k*k][cout][cin], then biases [cout]).

bf16=True rounds, as the device does, the weights, every stored activation
and every gradient that enters a matrix product to BF16 (round to nearest
even), so that device and oracle differ only in FP32 summation order.
bf16=False is the plain float64 network, pinned by central differences
(tests/test_oracle_unet.py).
"""
import numpy as np


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def layer_table(c0):
    c = [c0, 2 * c0, 4 * c0, 8 * c0]
    return [(3, c[0], 3, 1), (c[0], c[0], 3, 1), (c[0], c[1], 3, 1), (c[1], c[1], 3, 1), (c[1], c[2], 3, 1),
            (c[2], c[2], 3, 1), (c[2], c[3], 3, 1), (c[3], c[3], 3, 1), (c[3] + c[2], c[2], 3, 1),
            (c[2], c[2], 3, 1), (c[2] + c[1], c[1], 3, 1), (c[1], c[1], 3, 1), (c[1] + c[0], c[0], 3, 1),
            (c[0], c[0], 3, 1), (c[0], 2, 1, 0)]


def unpack(params, c0):
    out, off = [], 0
    for cin, cout, k, relu in layer_table(c0):
        nw = k * k * cout * cin
        W = params[off:off + nw].reshape(k * k, cout, cin)
        off += nw
        b = params[off:off + cout]
        off += cout
        out.append((W, b, k, relu))
    assert off == len(params)
    return out


def _offsets(k):
    return [(kk // 3 - 1, kk % 3 - 1) for kk in range(9)] if k == 3 else [(0, 0)]


def conv(X, W, b, k):
    """X [B,H,W,Cin] -> [B,H,W,Cout]: Y[p] = sum_kk X[p + o(kk)] W[kk]^T + b (periodic)."""
    Y = np.zeros(X.shape[:3] + (W.shape[1],), X.dtype)
    for kk, (dy, dx) in enumerate(_offsets(k)):
        Xs = np.roll(X, shift=(-dy, -dx), axis=(1, 2))
        Y += Xs @ W[kk].T.astype(X.dtype)
    return Y + b.astype(X.dtype)


def conv_bwd(X, W, k, gY):
    """gradients of conv w.r.t. W [kk][cout][cin], b and X."""
    dW = np.zeros(W.shape, np.float64)
    dX = np.zeros(X.shape, gY.dtype)
    for kk, (dy, dx) in enumerate(_offsets(k)):
        Xs = np.roll(X, shift=(-dy, -dx), axis=(1, 2))
        dW[kk] = np.einsum("bhwo,bhwi->oi", gY.astype(np.float64), Xs.astype(np.float64))
        dX += np.roll(gY @ W[kk].astype(gY.dtype), shift=(dy, dx), axis=(1, 2))
    db = gY.astype(np.float64).sum(axis=(0, 1, 2))
    return dW, db, dX


def pool(X):
    B, H, W, C = X.shape
    return X.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_bwd(g):
    return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2) * 0.25


def up(X):
    return np.repeat(np.repeat(X, 2, axis=1), 2, axis=2)


def up_bwd(g):
    B, H, W, C = g.shape
    return g.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def forward_backward(params, c0, x, target, mask, use_bf16=True, need_grad=True):
    """x [B,H,W,3], target [B,H,W,2], mask [B,H,W].  Returns (out, loss, grads)."""
    dt = np.float32 if use_bf16 else np.float64
    R = bf16 if use_bf16 else (lambda a: a)
    layers = unpack(np.asarray(params, np.float64), c0)
    Ws = [(R(W.astype(dt)), b.astype(dt), k, relu) for W, b, k, relu in layers]
    acts = {}

    def cv(i, X, relu_out=True):
        W, b, k, relu = Ws[i]
        Y = conv(X, W, b, k)
        if relu:
            Y = np.maximum(Y, 0)
        return R(Y.astype(dt)) if relu_out else Y.astype(dt)

    x0 = R(x.astype(dt))
    e0a = cv(0, x0); e0 = cv(1, e0a)
    p0 = R(pool(e0)); e1a = cv(2, p0); e1 = cv(3, e1a)
    p1 = R(pool(e1)); e2a = cv(4, p1); e2 = cv(5, e2a)
    p2 = R(pool(e2)); e3a = cv(6, p2); e3 = cv(7, e3a)
    c2 = np.concatenate([up(e3), e2], axis=3); d2a = cv(8, c2); d2 = cv(9, d2a)
    c1 = np.concatenate([up(d2), e1], axis=3); d1a = cv(10, c1); d1 = cv(11, d1a)
    c0t = np.concatenate([up(d1), e0], axis=3); d0a = cv(12, c0t); d0 = cv(13, d0a)
    out = cv(14, d0, relu_out=False)
    m = mask.astype(np.float64)
    denom = max(m.sum(), 1.0)
    diff = out.astype(np.float64) - target
    loss = float((m[..., None] * diff ** 2).sum() / (2 * denom))
    if not need_grad:
        return out, loss, None
    grads = [None] * len(Ws)

    def bw(i, X, Y, gY, need_dx=True):
        W, b, k, relu = Ws[i]
        g = gY.copy()
        if relu:
            g = np.where(Y > 0, g, 0)
        g = R(g.astype(dt))
        dW, db, dX = conv_bwd(X, W, k, g)
        grads[i] = (dW, db)
        return dX if need_dx else None

    gout = (m[..., None] * diff / denom).astype(dt)
    g_d0 = bw(14, d0, None, gout)
    g_d0a = bw(13, d0a, d0, g_d0)
    g_c0 = bw(12, c0t, d0a, g_d0a)
    C1 = d1.shape[3]
    g_d1 = up_bwd(g_c0[..., :C1]); g_e0 = g_c0[..., C1:].copy()
    g_d1a = bw(11, d1a, d1, g_d1)
    g_c1 = bw(10, c1, d1a, g_d1a)
    C1 = d2.shape[3]
    g_d2 = up_bwd(g_c1[..., :C1]); g_e1 = g_c1[..., C1:].copy()
    g_d2a = bw(9, d2a, d2, g_d2)
    g_c2 = bw(8, c2, d2a, g_d2a)
    C1 = e3.shape[3]
    g_e3 = up_bwd(g_c2[..., :C1]); g_e2 = g_c2[..., C1:].copy()
    g_e3a = bw(7, e3a, e3, g_e3)
    g_p2 = bw(6, p2, e3a, g_e3a)
    g_e2 = g_e2 + pool_bwd(g_p2)
    g_e2a = bw(5, e2a, e2, g_e2)
    g_p1 = bw(4, p1, e2a, g_e2a)
    g_e1 = g_e1 + pool_bwd(g_p1)
    g_e1a = bw(3, e1a, e1, g_e1)
    g_p0 = bw(2, p0, e1a, g_e1a)
    g_e0 = g_e0 + pool_bwd(g_p0)
    g_e0a = bw(1, e0a, e0, g_e0)
    bw(0, x0, e0a, g_e0a, need_dx=False)
    flat = np.concatenate([np.concatenate([dW.reshape(-1), db]) for dW, db in grads])
    return out, loss, flat


def param_count(c0):
    return sum(k * k * cout * cin + cout for cin, cout, k, _ in layer_table(c0))
