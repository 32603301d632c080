"""CPU restatement (numpy, FP64) of the learned directional guide -- test
infrastructure only, used to check the device implementation
(narvalengine_amd/csrc/guide.hip) and pinned itself by closed forms and
finite differences in tests/test_guide_ref.py.  The guide is a component of
this build with no counterpart in the reference, so there is no reference
file to follow; its definition is:

  encode(p): for each of 16 levels l with resolution res[l], the trilinear
    interpolation of 2 features at the 8 grid corners around p * res[l]
    (p clamped to [0,1], lower corner clamped to res-1); corner (x,y,z) is
    stored at x + (res+1) (y + (res+1) z) when (res+1)^3 <= T, else at
    (x ^ y*2654435761 ^ z*805459861) mod T (32-bit arithmetic).
  MLP: x0 = [encode(p) (32), direction (3), 0 x 13]; h_l = x_l W_l + b_l,
    x_{l+1} = ReLU(h_l) for l < 4; outputs = h_4[:40].  Device arithmetic:
    hash features, weights and layer inputs rounded to BF16, FP32
    accumulation; gradients reach the FP32 masters straight through.
  head: 8 lobes of (logit, mean xyz, log kappa): pi = softmax(logits),
    mu = mean / |mean|, kappa = exp(clamp(log kappa, -10, 8)).
  vMF: p(w) = kappa / (2 pi (1 - exp(-2 kappa))) exp(kappa (mu.w - 1)).
  loss = -sum_i r_i log q(w_i) / sum_i r_i,  q = sum_k pi_k vMF_k.
"""
from __future__ import annotations

import numpy as np

LEVELS, FEAT, LOBES, LAYERS = 16, 2, 8, 5
IN = [48, 64, 64, 64, 64]
OUT = [64, 64, 64, 64, 48]
W_OFF = np.cumsum([0] + [i * o for i, o in zip(IN, OUT)])
N_W = int(W_OFF[-1])
B_OFF = N_W + np.cumsum([0] + OUT)
N_PARAMS = int(B_OFF[-1])


def bf16(x):
    """Round to BF16 (nearest even) and back, as the device's f2bf."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFFFFFF
    return ((u.astype(np.uint32) >> 16) << 16).view(np.float32).astype(np.float64)


def unpack(mlp):
    W = [np.asarray(mlp[W_OFF[l]:W_OFF[l + 1]], np.float64).reshape(IN[l], OUT[l]) for l in range(LAYERS)]
    b = [np.asarray(mlp[B_OFF[l]:B_OFF[l + 1]], np.float64) for l in range(LAYERS)]
    return W, b


def corners(p, res, log2T):
    """(N, 8) table word offsets of feature 0 and (N, 8) weights for one level."""
    T = 1 << log2T
    p = np.clip(np.asarray(p, np.float64), 0.0, 1.0)
    s = (p * res).astype(np.float32).astype(np.float64)
    f = np.minimum(np.floor(s), res - 1)
    t = s - f
    fi = f.astype(np.int64)
    offs, ws = [], []
    r1 = res + 1
    for k in range(8):
        bits = np.array([k & 1, (k >> 1) & 1, (k >> 2) & 1])
        c = fi + bits
        w = np.prod(np.where(bits == 1, t, 1.0 - t), axis=1)
        if r1 ** 3 <= T:
            idx = c[:, 0] + r1 * (c[:, 1] + r1 * c[:, 2])
        else:
            x, y, z = (c[:, a].astype(np.uint64) for a in range(3))
            idx = ((x ^ ((y * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)) ^
                    ((z * np.uint64(805459861)) & np.uint64(0xFFFFFFFF))) & np.uint64(T - 1)).astype(np.int64)
        offs.append(idx)
        ws.append(w)
    return np.stack(offs, 1), np.stack(ws, 1)


def encode(p, table, res_list, log2T):
    T = 1 << log2T
    tab = np.asarray(table, np.float64).reshape(LEVELS, T, FEAT)
    out = np.zeros((len(p), LEVELS * FEAT))
    for l in range(LEVELS):
        idx, w = corners(p, res_list[l], log2T)
        for f in range(FEAT):
            out[:, 2 * l + f] = (w * tab[l, idx, f]).sum(1)
    return out


def forward(queries, table, mlp, res_list, log2T, quantize=True):
    """Raw outputs (N, 40) and the per-layer inputs (for backward)."""
    q = np.asarray(queries, np.float64)
    W, b = unpack(mlp)
    x = np.zeros((len(q), 48))
    x[:, :32] = encode(q[:, :3], bf16(table) if quantize else table, res_list, log2T)
    x[:, 32:35] = q[:, 3:6]
    if quantize:
        x = bf16(x)
        W = [bf16(w) for w in W]
    xs = [x]
    for l in range(LAYERS):
        h = xs[-1] @ W[l] + b[l]
        if l < LAYERS - 1:
            a = np.maximum(h, 0.0)
            xs.append(bf16(a) if quantize else a)
        else:
            out = h[:, :40]
    return out, xs


def head(y):
    y = np.asarray(y, np.float64).reshape(-1, LOBES, 5)
    lg = y[:, :, 0]
    pi = np.exp(lg - lg.max(1, keepdims=True))
    pi /= pi.sum(1, keepdims=True)
    m = y[:, :, 1:4]
    n = np.linalg.norm(m, axis=2, keepdims=True)
    mu = m / np.maximum(n, 1e-12)
    kap = np.exp(np.clip(y[:, :, 4], -10.0, 8.0))
    return pi, mu, kap


def vmf_pdf(kappa, cosv):
    kappa = np.asarray(kappa, np.float64)
    return kappa / (2 * np.pi * -np.expm1(-2 * kappa)) * np.exp(kappa * (np.asarray(cosv) - 1))


def mixture_pdf(y, w):
    pi, mu, kap = head(y)
    c = np.einsum("nkc,nc->nk", mu, np.asarray(w, np.float64))
    return (pi * vmf_pdf(kap, c)).sum(1)


def loss_and_grad(samples, table, mlp, res_list, log2T, quantize=False):
    """Weighted loss and its analytic gradient w.r.t. (table, mlp), FP64.
    quantize=True mirrors the device's BF16 points (straight-through)."""
    s = np.asarray(samples, np.float64)
    pos, w, r = s[:, 0:3], s[:, 6:9], s[:, 9]
    y, xs = forward(s[:, :6], table, mlp, res_list, log2T, quantize)
    pi, mu, kap = head(y)
    c = np.einsum("nkc,nc->nk", mu, w)
    v = vmf_pdf(kap, c)
    q = np.maximum((pi * v).sum(1), 1e-30)
    R = r.sum()
    loss = float((-r * np.log(q)).sum() / R)
    sc = (-r / q)[:, None]
    yy = y.reshape(-1, LOBES, 5)
    nm = np.maximum(np.linalg.norm(yy[:, :, 1:4], axis=2), 1e-12)
    dy = np.zeros((len(s), LOBES, 5))
    dy[:, :, 0] = sc * pi * (v - q[:, None])
    pv = sc * pi * v
    dy[:, :, 1:4] = (pv * kap / nm)[:, :, None] * (w[:, None, :] - c[:, :, None] * mu)
    dlog = 1 / kap - 2 / np.expm1(2 * kap) + (c - 1)  # d log vMF / d kappa
    dy[:, :, 4] = np.where((yy[:, :, 4] < 8.0) & (yy[:, :, 4] > -10.0), pv * kap * dlog, 0.0)
    g = np.zeros((len(s), 48))
    g[:, :40] = dy.reshape(len(s), 40)
    W, _ = unpack(mlp)
    if quantize:
        W = [bf16(x) for x in W]
    mlp_grad = np.zeros(N_PARAMS)
    for l in reversed(range(LAYERS)):
        gq = bf16(g) if quantize else g
        mlp_grad[W_OFF[l]:W_OFF[l + 1]] = (xs[l].T @ gq).reshape(-1) / R
        mlp_grad[B_OFF[l]:B_OFF[l + 1]] = gq.sum(0) / R
        gx = gq @ W[l].T
        if l > 0:
            g = np.where(xs[l] > 0, gx, 0.0)
        else:
            denc = gx[:, :32]
    T = 1 << log2T
    table_grad = np.zeros((LEVELS, T, FEAT))
    for l in range(LEVELS):
        idx, wt = corners(pos, res_list[l], log2T)
        for f in range(FEAT):
            np.add.at(table_grad[l, :, f], idx.reshape(-1), (wt * denc[:, 2 * l + f][:, None]).reshape(-1) / R)
    return loss, table_grad.reshape(-1), mlp_grad
