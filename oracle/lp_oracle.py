"""CPU oracle of the learned proposal's network (test infrastructure only).

A numpy restatement, in float64, of the two-tower encoder and its InfoNCE
training gradient as the library computes them: every matrix product takes
operands rounded to bfloat16 (round to nearest even) and accumulates in
full precision, biases and the softmax stay in floating point. With
`bf16=False` the same code is the exact float64 model, which the tests pin
against central finite differences.

Parameter layout (one flat vector): query tower then key tower, each
W1 [240][256], b1, W2 [256][256], b2, W3 [256][256], b3, W4 [256][128], b4.
The learned proposal has no counterpart in the reference; nothing here
restates reference code.
"""
import numpy as np

DIN = (240, 256, 256, 256)
DOUT = (256, 256, 256, 128)


def _offsets():
    offs, o = [], 0
    for l in range(4):
        w = o
        o += DIN[l] * DOUT[l]
        b = o
        o += DOUT[l]
        offs.append((w, b))
    return offs, o


LAYERS, TOWER = _offsets()


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64).reshape(np.shape(x))


def _mm(a, b, rnd):
    if rnd:
        return bf16(a) @ bf16(b)
    return np.asarray(a, np.float64) @ np.asarray(b, np.float64)


def tower_params(p, t):
    base = t * TOWER
    out = []
    for l in range(4):
        w, b = LAYERS[l]
        W = np.asarray(p[base + w:base + w + DIN[l] * DOUT[l]], np.float64).reshape(DIN[l], DOUT[l])
        bb = np.asarray(p[base + b:base + b + DOUT[l]], np.float64)
        out.append((W, bb))
    return out


def forward(p, t, X, rnd=True):
    """Activations [X, H1, H2, H3, E] of tower t."""
    acts = [np.asarray(X, np.float64)]
    h = acts[0]
    for l, (W, b) in enumerate(tower_params(p, t)):
        h = _mm(h, W, rnd) + b
        if l < 3:
            h = np.maximum(h, 0.0)
        acts.append(h)
    return acts


def backward(p, t, acts, dE, rnd=True):
    """Gradient of tower t's parameters (flat, TOWER floats) from dE."""
    g = np.zeros(TOWER)
    params = tower_params(p, t)
    d = np.asarray(dE, np.float64)
    for l in range(3, -1, -1):
        W, _ = params[l]
        inp = acts[l]
        w, b = LAYERS[l]
        g[w:w + DIN[l] * DOUT[l]] = _mm(inp.T, d, rnd).reshape(-1)
        g[b:b + DOUT[l]] = d.sum(0)
        if l > 0:
            d = _mm(d, W.T, rnd) * (acts[l] > 0)
    return g


def loss_grad(p, qfeat, kfeat, inv_tau, rnd=True):
    """Mean InfoNCE loss (row b's positive is key b) and the full gradient."""
    qa = forward(p, 0, qfeat, rnd)
    ka = forward(p, 1, kfeat, rnd)
    Q, K = qa[-1], ka[-1]
    B = Q.shape[0]
    G = _mm(Q, K.T, rnd)
    z = G * inv_tau
    z = z - z.max(1, keepdims=True)
    P = np.exp(z)
    P /= P.sum(1, keepdims=True)
    loss = float(np.mean(-np.log(P[np.arange(B), np.arange(B)])))
    Y = np.zeros_like(P)
    Y[np.arange(B), np.arange(B)] = 1.0
    dG = inv_tau * (P - Y) / B
    dQ = _mm(dG, K, rnd)
    dK = _mm(dG.T, Q, rnd)
    grad = np.concatenate([backward(p, 0, qa, dQ, rnd), backward(p, 1, ka, dK, rnd)])
    return loss, grad


def features_from_neighbours(offsets, values, canon, mu, sigma):
    """Query features: the neighbour at canonical offset canon[i] fills slot
    i with (dx/16, dy/16, (x - mu)/sigma); neighbours off the stencil and
    empty (NaN) entries leave their slots zero."""
    offsets = np.asarray(offsets, np.int64).reshape(-1, 80, 2)
    values = np.asarray(values, np.float64).reshape(-1, 80)
    slot = {(int(a), int(b)): i for i, (a, b) in enumerate(np.asarray(canon).reshape(80, 2))}
    f = np.zeros((values.shape[0], 80, 3))
    for r in range(values.shape[0]):
        for j in range(80):
            if np.isnan(values[r, j]):
                continue
            s = slot.get((int(offsets[r, j, 0]), int(offsets[r, j, 1])))
            if s is None:
                continue
            f[r, s] = (offsets[r, j, 0] / 16.0, offsets[r, j, 1] / 16.0, (values[r, j] - mu) / sigma)
    return f.reshape(-1, 240)


def key_features(ti, dims, canon, positions, mu, sigma):
    """Key features of TI positions on the canonical offsets (zero slots
    outside the TI)."""
    s0, s1 = dims
    img = np.asarray(ti, np.float64).reshape(s1, s0)
    pos = np.asarray(positions, np.int64)
    x0, x1 = pos % s0, pos // s0
    f = np.zeros((pos.size, 80, 3))
    for i, (d0, d1) in enumerate(np.asarray(canon).reshape(80, 2)):
        t0, t1 = x0 + d0, x1 + d1
        ok = (t0 >= 0) & (t0 < s0) & (t1 >= 0) & (t1 < s1)
        f[ok, i, 0] = d0 / 16.0
        f[ok, i, 1] = d1 / 16.0
        f[ok, i, 2] = (img[t1[ok], t0[ok]] - mu) / sigma
    return f.reshape(-1, 240)
