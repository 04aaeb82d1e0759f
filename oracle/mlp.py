"""numpy restatement of the learned-guess network (test infrastructure only).

Same architecture, parameter layout and bf16 rounding points as the HIP
kernels: every GEMM operand (activations, gradients, weights) is rounded to
bf16 before the product, products are summed in float64 here (fp32 on the
GPU), biases, activations and bias gradients stay unrounded.  With
round_bf16=False the model is smooth and its analytic gradient can be pinned
by finite differences.
"""
import numpy as np

LAYOUT = {}
_o = 0
for _name, _size in (("w0", 64 * 32), ("b0", 64), ("w1", 64 * 128), ("b1", 64), ("w2", 64 * 128),
                     ("b2", 64), ("w3", 16 * 64), ("b3", 16)):
    LAYOUT[_name] = (_o, _size)
    _o += _size
NPARAMS = _o


def bf16(a):
    """round-to-nearest-even of float32 values to bf16, returned as float64"""
    a32 = np.asarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    g = lambda n: p[LAYOUT[n][0]:LAYOUT[n][0] + LAYOUT[n][1]]
    return dict(w0=g("w0").reshape(64, 32), b0=g("b0"), w1=g("w1").reshape(64, 128), b1=g("b1"),
                w2=g("w2").reshape(64, 128), b2=g("b2"), w3=g("w3").reshape(16, 64), b3=g("b3"))


def agg(h, off, idx):
    deg = np.diff(off)
    out = np.zeros_like(h)
    rows = np.repeat(np.arange(len(deg)), deg)
    np.add.at(out, rows, h[idx])
    nz = deg > 0
    out[nz] /= deg[nz, None]
    return out


def agg_t(d, off, idx):
    """transpose of agg: dh[n] += sum_{c: n in N(c)} d[c] / deg(c)"""
    deg = np.diff(off)
    rows = np.repeat(np.arange(len(deg)), deg)
    out = np.zeros_like(d)
    w = d / np.maximum(deg, 1)[:, None]
    np.add.at(out, idx, w[rows])
    return out


def forward(p, x, off, idx, round_bf16=True):
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    P = unpack(p)
    W = {k: r(v) if k.startswith("w") else v for k, v in P.items()}
    x = np.asarray(x, np.float64)
    h0 = np.maximum(r(x) @ W["w0"].T + P["b0"], 0)
    a1 = agg(h0, off, idx)
    c1 = np.concatenate([h0, a1], 1)
    h1 = np.maximum(r(c1) @ W["w1"].T + P["b1"], 0)
    a2 = agg(h1, off, idx)
    c2 = np.concatenate([h1, a2], 1)
    h2 = np.maximum(r(c2) @ W["w2"].T + P["b2"], 0)
    out = r(h2) @ W["w3"].T + P["b3"]
    return out, dict(x=x, h0=h0, a1=a1, c1=c1, h1=h1, a2=a2, c2=c2, h2=h2, W=W)


def loss_and_grad(p, x, y, off, idx, round_bf16=True):
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    out, t = forward(p, x, off, idx, round_bf16)
    W = t["W"]
    N = x.shape[0]
    e = out[:, :5] - np.asarray(y, np.float64)[:, :5]
    loss = float((e * e).sum() / (5 * N))
    dout = np.zeros((N, 16))
    dout[:, :5] = 2 * e / (5 * N)
    g = {}
    g["w3"] = r(dout).T @ r(t["h2"])
    g["b3"] = dout.sum(0)
    g2 = (r(dout) @ W["w3"]) * (t["h2"] > 0)
    g["w2"] = r(g2).T @ r(t["c2"])
    g["b2"] = g2.sum(0)
    dc2 = r(g2) @ W["w2"]
    g1 = (dc2[:, :64] + agg_t(dc2[:, 64:], off, idx)) * (t["h1"] > 0)
    g["w1"] = r(g1).T @ r(t["c1"])
    g["b1"] = g1.sum(0)
    dc1 = r(g1) @ W["w1"]
    g0 = (dc1[:, :64] + agg_t(dc1[:, 64:], off, idx)) * (t["h0"] > 0)
    g["w0"] = r(g0).T @ r(t["x"])
    g["b0"] = g0.sum(0)
    g["w3"][5:] = 0
    g["b3"][5:] = 0
    grad = np.zeros(NPARAMS)
    for k, (o, s) in LAYOUT.items():
        grad[o:o + s] = g[k].ravel()
    return loss, grad, out
