"""FP64 restatement of the learned predictor's network (test infrastructure only).

3-level 3D U-Net on an R^3 grid, channels-last, built from the oracle's C
convolutions: c1a c1b | pool | c2a c2b | pool | c3a c3b | up, concat c2b | d2a d2b |
up, concat c1b | d1a d1b | out (1x1x1).  ReLU after every 3x3x3 convolution, 2x
average pooling, nearest 2x upsampling, loss 0.5*mean((y[:, :4] - t)^2).  Parameter
vector layout: per layer W[tap][cin][cout] then b[cout].

`q` (optional) is a rounding function applied where the GPU network stores
BF16: input features, weights, every hidden activation, pooled maps and the
masked output gradient that feeds each layer's backward convolutions.  With
q=None the network is plain FP64 (pinned by finite differences in the tests);
with q=bf16_round it models the GPU's storage precision so that the GPU's
gradients can be compared at FP32-accumulation tolerance.
"""
import numpy as np

import oracle as O

SPEC = [  # cin, cout, ntap, relu, level
    (16, 16, 27, 1, 0), (16, 16, 27, 1, 0), (16, 32, 27, 1, 1), (32, 32, 27, 1, 1),
    (32, 64, 27, 1, 2), (64, 64, 27, 1, 2), (96, 32, 27, 1, 1), (32, 32, 27, 1, 1),
    (48, 16, 27, 1, 0), (16, 16, 27, 1, 0), (16, 16, 1, 0, 0),
]
NOUT = 4


def bf16_round(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(a):
    return a


def unpack(params):
    out, o = [], 0
    for cin, cout, nt, _, _ in SPEC:
        w = params[o:o + nt * cin * cout]
        o += nt * cin * cout
        b = params[o:o + cout]
        o += cout
        out.append((np.asarray(w, np.float64), np.asarray(b, np.float64)))
    return out


def pool(x, R, C):
    return x.reshape(R // 2, 2, R // 2, 2, R // 2, 2, C).mean(axis=(1, 3, 5)).reshape(-1)


def pool_bwd(g, R, C):
    g = g.reshape(R // 2, R // 2, R // 2, C) / 8.0
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2).reshape(-1)


def up(x, Rc, C):
    g = x.reshape(Rc, Rc, Rc, C)
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2).reshape(-1)


def up_bwd(g, Rf, C):
    return g.reshape(Rf // 2, 2, Rf // 2, 2, Rf // 2, 2, C).sum(axis=(1, 3, 5)).reshape(-1)


def cat(a, Ca, b, Cb, nv):
    return np.concatenate([a.reshape(nv, Ca), b.reshape(nv, Cb)], axis=1).reshape(-1)


def split(g, Ca, Cb, nv):
    g = g.reshape(nv, Ca + Cb)
    return g[:, :Ca].reshape(-1).copy(), g[:, Ca:].reshape(-1).copy()


def forward(params, feat, R, q=None):
    """feat: R^3 x 4.  Returns (output R^3 x 16, cache)."""
    q = q or _ident
    P = [(q(w), b) for w, b in unpack(params)]
    n0, n1, n2 = R ** 3, (R // 2) ** 3, (R // 4) ** 3
    R2, R4 = R // 2, R // 4
    x0 = np.zeros((n0, 16))
    x0[:, :4] = np.asarray(feat, np.float64).reshape(n0, 4)
    x0 = q(x0.reshape(-1))
    cv = lambda l, x, Rl: q(O.conv3d(Rl, SPEC[l][2], SPEC[l][0], SPEC[l][1], x, P[l][0], P[l][1], bool(SPEC[l][3])))
    c = {}
    c['x0'] = x0
    c['e1a'] = cv(0, x0, R)
    c['e1b'] = cv(1, c['e1a'], R)
    c['p1'] = q(pool(c['e1b'], R, 16))
    c['e2a'] = cv(2, c['p1'], R2)
    c['e2b'] = cv(3, c['e2a'], R2)
    c['p2'] = q(pool(c['e2b'], R2, 32))
    c['e3a'] = cv(4, c['p2'], R4)
    c['e3b'] = cv(5, c['e3a'], R4)
    c['d2in'] = cat(up(c['e3b'], R4, 64), 64, c['e2b'], 32, n1)
    c['d2a'] = cv(6, c['d2in'], R2)
    c['d2b'] = cv(7, c['d2a'], R2)
    c['d1in'] = cat(up(c['d2b'], R2, 32), 32, c['e1b'], 16, n0)
    c['d1a'] = cv(8, c['d1in'], R)
    c['d1b'] = cv(9, c['d1a'], R)
    c['out'] = O.conv3d(R, 1, 16, 16, c['d1b'], P[10][0], P[10][1], False)
    return c['out'], c


def loss_and_grad(params, feat, tgt, R, q=None):
    """0.5*mean((out[:, :4]-tgt)^2) and its gradient w.r.t. every parameter."""
    out, c = forward(params, feat, R, q)
    q = q or _ident
    P = [(q(w), b) for w, b in unpack(params)]
    n0 = R ** 3
    R2, R4 = R // 2, R // 4
    y = out.reshape(n0, 16)
    d = y[:, :NOUT] - np.asarray(tgt, np.float64).reshape(n0, NOUT)
    L = 0.5 * np.mean(d * d)
    g = np.zeros((n0, 16))
    g[:, :NOUT] = d / d.size
    grads = [None] * len(SPEC)

    def layer_bwd(l, x, yv, gy, Rl):
        cin, cout, nt, relu, _ = SPEC[l]
        gz = q(gy * (yv > 0) if relu else gy)
        gw, gb = O.conv3d_bwd_weight(Rl, nt, cin, cout, x, gz)
        grads[l] = (gw, gb)
        return O.conv3d_bwd_data(Rl, nt, cin, cout, gz, P[l][0])

    gd1b = layer_bwd(10, c['d1b'], c['out'], g.reshape(-1), R)
    gd1a = layer_bwd(9, c['d1a'], c['d1b'], gd1b, R)
    gd1in = layer_bwd(8, c['d1in'], c['d1a'], gd1a, R)
    gu1, ge1b = split(gd1in, 32, 16, n0)
    gd2b = up_bwd(gu1, R, 32)
    gd2a = layer_bwd(7, c['d2a'], c['d2b'], gd2b, R2)
    gd2in = layer_bwd(6, c['d2in'], c['d2a'], gd2a, R2)
    gu2, ge2b = split(gd2in, 64, 32, (R2) ** 3)
    ge3b = up_bwd(gu2, R2, 64)
    ge3a = layer_bwd(5, c['e3a'], c['e3b'], ge3b, R4)
    gp2 = layer_bwd(4, c['p2'], c['e3a'], ge3a, R4)
    ge2b = ge2b + pool_bwd(gp2, R2, 32)
    ge2a = layer_bwd(3, c['e2a'], c['e2b'], ge2b, R2)
    gp1 = layer_bwd(2, c['p1'], c['e2a'], ge2a, R2)
    ge1b = ge1b + pool_bwd(gp1, R, 16)
    ge1a = layer_bwd(1, c['e1a'], c['e1b'], ge1b, R)
    layer_bwd(0, c['x0'], c['e1a'], ge1a, R)
    flat = np.concatenate([np.concatenate([gw, gb]) for gw, gb in grads])
    return L, flat, out
