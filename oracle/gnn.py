"""CPU oracle of the active-set predictor (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

The predictor has no counterpart in the reference; this file states its
definition once, independently of the HIP kernels: the per-vertex pack of
quantised values, the 16 input features, the two layers, the weighted
cross-entropy and its hand-derived gradient.  Pinned by tests/test_gnn_oracle.py:
the quantisers against Python integer arithmetic and torch's bfloat16
rounding, the features against values worked out by hand on a small graph,
and the gradient against finite differences and torch autograd.

Graphs are CSR arrays (rp int64, col int32, optional w), symmetric.
"""
import numpy as np

GNN_IN, GNN_HID = 16, 32
GP_W1, GP_B1 = 0, 512
GP_W2S, GP_W2N, GP_B2 = 544, 576, 608
GNN_P = 609
NO_MOVE = -3
FLT_MAX = np.float32(3.4028234663852886e38)


def bf16_rne(x):
    """float32 -> bfloat16 bits, round to nearest even."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = u + 0x7FFF + ((u >> 16) & 1)
    return ((u >> 16) & 0xFFFF).astype(np.uint16)


def bf16_f32(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def logq(v):
    """8*floor(log2 v) + the three bits below the leading one, v >= 1."""
    v = np.asarray(v, dtype=np.uint64)
    out = np.empty(v.shape, np.int64)
    flat, of = v.reshape(-1), out.reshape(-1)
    for k, x in enumerate(flat.tolist()):
        e = int(x).bit_length() - 1
        f = (x >> (e - 3)) & 7 if e >= 3 else (x << (3 - e)) & 7
        of[k] = 8 * e + f
    return out


def margin_q(gain, deg, first):
    """gain per (degree+1) in 1/64 steps, clamped to [-127, 127]; -128 for a
    cached NO_MOVE gain (-FLT_MAX); 0 before the first sweep scored it."""
    gain = np.asarray(gain, np.float32)
    if first:
        return np.zeros(gain.shape, np.int64)
    with np.errstate(over="ignore", invalid="ignore"):
        m = np.rint(gain / (np.asarray(deg, np.int64) + 1).astype(np.float32) * np.float32(64.0))
    m = np.clip(m, -127, 127).astype(np.int64)
    return np.where(gain > -FLT_MAX, m, -128)


def pack(rp, part, dest, gain, stamp, cur, first):
    deg = np.diff(rp).astype(np.int64)
    return {
        "part": np.asarray(part, np.int64),
        "lq": np.minimum(logq(deg + 1), 255),
        "mq": margin_q(gain, deg, first),
        "moved": (np.asarray(stamp) == cur - 1).astype(np.int64),
        "act": (np.asarray(dest) != NO_MOVE).astype(np.int64),
    }


def features(rp, col, w, vw, part, dest, gain, stamp, cur, first, vol):
    """Input rows as float32 before bf16 rounding, shape (n, 16)."""
    rp = np.asarray(rp, np.int64)
    col = np.asarray(col, np.int64)
    n = len(rp) - 1
    pk = pack(rp, part, dest, gain, stamp, cur, first)
    deg = np.diff(rp)
    row = np.repeat(np.arange(n), deg)
    wt = np.ones(len(col), np.int64) if w is None else np.asarray(w, np.int64)
    same = (pk["part"][col] == pk["part"][row]).astype(np.int64)
    terms = [pk["lq"][col], pk["mq"][col], pk["moved"][col], pk["act"][col], same,
             (pk["mq"][col] > pk["mq"][row]).astype(np.int64), same & pk["act"][col], same & pk["moved"][col],
             wt, same * wt]
    S = np.zeros((10, n), np.int64)
    for k, t in enumerate(terms):
        np.add.at(S[k], row, t)
    f32 = np.float32
    x = np.zeros((n, GNN_IN), np.float32)
    vw = np.asarray(vw, np.int64)
    volp = np.asarray(vol, np.int64)[pk["part"]]
    x[:, 0] = pk["lq"].astype(f32) / f32(256)
    x[:, 1] = logq(np.where(vw > 0, vw + 1, 1)).astype(f32) / f32(512)
    x[:, 2] = pk["mq"].astype(f32) / f32(128)
    x[:, 3] = pk["moved"].astype(f32)
    x[:, 4] = pk["act"].astype(f32)
    with np.errstate(invalid="ignore", divide="ignore"):
        x[:, 5] = np.where(S[8] > 0, S[9].astype(f32) / S[8].astype(f32), f32(0))
    x[:, 6] = logq(np.where(volp > 0, volp + 1, 1)).astype(f32) / f32(512)
    x[:, 7] = 1
    D = np.maximum(deg, 1).astype(f32)
    x[:, 8] = S[0].astype(f32) / (D * f32(256))
    x[:, 9] = S[1].astype(f32) / (D * f32(128))
    for k in range(2, 8):
        x[:, 8 + k] = S[k].astype(f32) / D
    return x


def rows(*args, **kw):
    """bf16 bits of the input rows, shape (n, 16)."""
    return bf16_rne(features(*args, **kw))


def _sum_nbr(rp, col, v):
    """sum of v over the neighbours of every vertex (float64)"""
    rp = np.asarray(rp, np.int64)
    n = len(rp) - 1
    row = np.repeat(np.arange(n), np.diff(rp))
    acc = np.zeros(n, np.float64)
    np.add.at(acc, row, np.asarray(v, np.float64)[np.asarray(col, np.int64)])
    return acc


def _mean_nbr(rp, col, v):
    return _sum_nbr(rp, col, v) / np.maximum(np.diff(np.asarray(rp, np.int64)), 1)


def forward(rp, col, rows_bits, params, round_hidden=True, round_w=True):
    """Returns (in, z, h, s, t, logit) in float64; logit for every vertex.
    The rounding flags switch off the bf16 steps (for derivative checks)."""
    p = np.asarray(params, np.float64 if not round_w else np.float32)
    IN = bf16_f32(rows_bits).astype(np.float64)
    if round_w:
        W1 = bf16_f32(bf16_rne(p[GP_W1:GP_B1])).astype(np.float64).reshape(GNN_IN, GNN_HID)
    else:
        W1 = p[GP_W1:GP_B1].reshape(GNN_IN, GNN_HID)
    z = IN @ W1 + p[GP_B1:GP_W2S].astype(np.float64)
    h = np.maximum(z, 0.0)
    if round_hidden:
        h = bf16_f32(bf16_rne(h.astype(np.float32))).astype(np.float64)
    s = h @ p[GP_W2N:GP_B2].astype(np.float64)
    t = h @ p[GP_W2S:GP_W2N].astype(np.float64) + float(p[GP_B2])
    logit = t + _mean_nbr(rp, col, s)
    return IN, z, h, s, t, logit


def softplus(x):
    return np.maximum(x, 0) + np.log1p(np.exp(-np.abs(x)))


def loss_and_grad(rp, col, rows_bits, params, active, y, pos_weight, exact=False):
    """Mean weighted BCE over active vertices and its gradient (GNN_P).
    exact=True drops every bf16 rounding (W1, h, dZ): the true derivative of
    the smooth model, for derivative checks."""
    p = np.asarray(params, np.float64 if exact else np.float32)
    IN, z, h, s, t, L = forward(rp, col, rows_bits, p, round_hidden=not exact, round_w=not exact)
    act = np.asarray(active).astype(bool)
    y = np.asarray(y).astype(np.float64)
    na = max(int(act.sum()), 1)
    pw = float(pos_weight)
    loss = float(np.sum(np.where(act, pw * y * softplus(-L) + (1 - y) * softplus(L), 0.0)) / na)
    sg = 1.0 / (1.0 + np.exp(-L))
    g = np.where(act, (sg * (pw * y + 1 - y) - pw * y) / na, 0.0)
    deg = np.diff(np.asarray(rp, np.int64))
    r = g / np.maximum(deg, 1)
    ds = _sum_nbr(rp, col, r)  # adjoint of the neighbour mean (symmetric graph)
    w2s = p[GP_W2S:GP_W2N].astype(np.float64)
    w2n = p[GP_W2N:GP_B2].astype(np.float64)
    dz = (g[:, None] * w2s[None, :] + ds[:, None] * w2n[None, :]) * (h > 0)
    dzm = dz if exact else bf16_f32(bf16_rne(dz.astype(np.float32))).astype(np.float64)
    grad = np.zeros(GNN_P, np.float64)
    grad[GP_W1:GP_B1] = (IN.T @ dzm).reshape(-1)
    grad[GP_B1:GP_W2S] = dz.sum(0)
    grad[GP_W2S:GP_W2N] = (g[:, None] * h).sum(0)
    grad[GP_W2N:GP_B2] = (ds[:, None] * h).sum(0)
    grad[GP_B2] = g.sum()
    return loss, grad, L
