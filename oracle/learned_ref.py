"""FP64 CPU oracle of the learned sparse-view FBP (test infrastructure only).

No reference code exists for this part (it extends the reference's Radon/FBP
operator, src/kernelRADN.h:212-259, with a learned filter and a residual
CNN).  This is a float64 numpy restatement of the model the GPU kernels
compute, forward and backward:

  r0  = BP(h * p),  h[m] = taps[|m|]                       (std_fbp geometry)
  a1  = bf(relu(conv(bf(r0); bf(W1)) + b1))               1 -> C
  a2  = bf(relu(conv(a1; bf(W2)) + b2))                   C -> C
  a3  = bf(relu(conv(a2; bf(W3)) + b3))                   C -> C
  out = r0 + conv(a3; bf(W4)) + b4                        C -> 1

with C = 16 or 32 hidden channels (keyword C, default 16).
  loss = mean((out - target)^2)

bf() rounds to bfloat16 (nearest even) when `bf16=True`, the identity
otherwise; gradients pass straight through bf() (the GPU differentiates the
same way).  3x3 convolutions with zero padding, tap t = 3*ky + kx reading
pixel (i+ky-1, j+kx-1).  Parameter vector layout: W1[C][9] b1[C]
W2[C][9][C] b2[C] W3[C][9][C] b3[C] W4[9][C] b4 taps[n_det].

Pinned by tests/test_oracle_learned.py: central finite differences of the
FP64 loss (bf16 off) for every parameter block and for r0, the adjoint
identity <BP q, r> = <q, BP^T r>, and the filter gradient against the
convolution it differentiates.
"""
import numpy as np

from . import std_fbp

def layout(C):
    """Parameter layout for C hidden channels: ({block: (offset, size)},
    offset of the taps)."""
    off = {}
    o = 0
    for name, size in (("W1", C * 9), ("b1", C), ("W2", C * 9 * C), ("b2", C), ("W3", C * 9 * C), ("b3", C),
                       ("W4", 9 * C), ("b4", 1)):
        off[name] = (o, size)
        o += size
    return off, o


C = 16
OFF, TAPS = layout(C)  # the 16-channel layout (TAPS = 4945)


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(prm, C=16):
    prm = np.asarray(prm, np.float64)
    off, taps = layout(C)
    g = {}
    for k, (o, s) in off.items():
        g[k] = prm[o:o + s]
    g["W1"] = g["W1"].reshape(C, 9)
    g["W2"] = g["W2"].reshape(C, 9, C)
    g["W3"] = g["W3"].reshape(C, 9, C)
    g["W4"] = g["W4"].reshape(9, C)
    g["taps"] = prm[taps:]
    return g


def pack(g, n_det, C=16):
    off, taps = layout(C)
    out = np.zeros(taps + n_det)
    for k, (o, s) in off.items():
        out[o:o + s] = np.asarray(g[k]).reshape(-1)
    out[taps:] = g["taps"]
    return out


def conv3x3(x, W):
    """x (B, n, n, ci), W (co, 9, ci) -> (B, n, n, co), zero padding."""
    B, n, _, ci = x.shape
    xp = np.zeros((B, n + 2, n + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    z = 0.0
    for t in range(9):
        ky, kx = divmod(t, 3)
        z = z + xp[:, ky:ky + n, kx:kx + n, :] @ W[:, t, :].T
    return z


def conv3x3_adjoint_input(dz, W):
    """Adjoint of conv3x3 with respect to x: dx[q] = sum_t W_t^T dz[q - off(t)]."""
    B, n, _, co = dz.shape
    ci = W.shape[2]
    dx = np.zeros((B, n + 2, n + 2, ci))
    for t in range(9):
        ky, kx = divmod(t, 3)
        dx[:, ky:ky + n, kx:kx + n, :] += dz @ W[:, t, :]
    return dx[:, 1:-1, 1:-1]


def conv3x3_weight_grad(dz, x):
    """dW[co][t][ci] = sum_p dz[p][co] x[p + off(t)][ci]."""
    B, n, _, ci = x.shape
    co = dz.shape[3]
    xp = np.zeros((B, n + 2, n + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    dW = np.zeros((co, 9, ci))
    dzf = dz.reshape(-1, co)
    for t in range(9):
        ky, kx = divmod(t, 3)
        dW[:, t, :] = dzf.T @ xp[:, ky:ky + n, kx:kx + n, :].reshape(-1, ci)
    return dW


def cnn_forward(prm, r0, bf=True, C=16):
    """r0 (B, n, n) -> dict with the activations and `out`."""
    g = unpack(prm, C)
    q = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    r0 = np.asarray(r0, np.float64)
    x0 = q(r0)[..., None]
    z1 = conv3x3(x0, q(g["W1"])[:, :, None]) + g["b1"]
    a1 = q(np.maximum(z1, 0.0))
    z2 = conv3x3(a1, q(g["W2"])) + g["b2"]
    a2 = q(np.maximum(z2, 0.0))
    z3 = conv3x3(a2, q(g["W3"])) + g["b3"]
    a3 = q(np.maximum(z3, 0.0))
    out = r0 + conv3x3(a3, q(g["W4"])[None, :, :])[..., 0] + g["b4"][0]
    return {"r0": r0, "x0": x0, "a1": a1, "a2": a2, "a3": a3, "out": out}


def cnn_backward(prm, cache, dout, bf=True, C=16):
    """Gradients of a scalar loss with d loss/d out = dout (B, n, n): the
    parameter gradient (taps part zero) and d loss / d r0."""
    g = unpack(prm, C)
    q = bf16 if bf else (lambda v: np.asarray(v, np.float64))
    W1, W2, W3, W4 = q(g["W1"])[:, :, None], q(g["W2"]), q(g["W3"]), q(g["W4"])[None, :, :]
    d = dout[..., None]
    grads = {"W4": conv3x3_weight_grad(d, cache["a3"])[0], "b4": np.array([dout.sum()])}
    dz3 = conv3x3_adjoint_input(d, W4) * (cache["a3"] > 0)
    grads["W3"] = conv3x3_weight_grad(dz3, cache["a2"])
    grads["b3"] = dz3.sum(axis=(0, 1, 2))
    dz2 = conv3x3_adjoint_input(dz3, W3) * (cache["a2"] > 0)
    grads["W2"] = conv3x3_weight_grad(dz2, cache["a1"])
    grads["b2"] = dz2.sum(axis=(0, 1, 2))
    dz1 = conv3x3_adjoint_input(dz2, W2) * (cache["a1"] > 0)
    grads["W1"] = conv3x3_weight_grad(dz1, cache["x0"])[:, :, 0]
    grads["b1"] = dz1.sum(axis=(0, 1, 2))
    dr0 = dout + conv3x3_adjoint_input(dz1, W1)[..., 0]
    grads["dz1"], grads["dz2"], grads["dz3"] = dz1, dz2, dz3
    return grads, dr0


def full_taps(sym):
    sym = np.asarray(sym, np.float64)
    n_det = sym.size
    return sym[np.abs(np.arange(-(n_det - 1), n_det))]


def backproject_adjoint(dr, n_angles, n_det, angles=None):
    """BP^T of std_fbp.backproject: dq[a][k] = (pi/nang) * sum over pixels of
    the pixel's linear-interpolation weight on bin k times dr."""
    dr = np.asarray(dr, np.float64)
    n = dr.shape[0]
    delta = 2.0 / n
    det_c = 0.5 * (n_det - 1)
    cs, sn = std_fbp.angle_table(n_angles, angles)
    x = (np.arange(n) + 0.5) * delta - 1.0
    y = 1.0 - (np.arange(n) + 0.5) * delta
    X, Y = x[None, :], y[:, None]
    dq = np.zeros((n_angles, n_det))
    for a in range(n_angles):
        kf = (X * cs[a] + Y * sn[a]) / delta + det_c
        k0 = np.floor(kf).astype(np.int64)
        w = kf - k0
        for k, wt in ((k0, 1.0 - w), (k0 + 1, w)):
            ok = (k >= 0) & (k < n_det)
            np.add.at(dq[a], k[ok], (wt * dr)[ok])
    return dq * (np.pi / n_angles)


def filter_grad(dq, sino):
    """d loss / d full taps (index m + n_det - 1) of q = filter_rows(p, h):
    dh[m] = sum over rows and k of dq[k] p[k - m]."""
    dq = np.asarray(dq, np.float64).reshape(-1, dq.shape[-1])
    p = np.asarray(sino, np.float64).reshape(-1, dq.shape[-1])
    n_det = dq.shape[-1]
    dh = np.zeros(2 * n_det - 1)
    for r in range(dq.shape[0]):
        dh += np.correlate(dq[r], p[r], mode="full")  # index m + n_det - 1
    return dh


def fold_taps(dfull):
    n_det = (dfull.size + 1) // 2
    c = n_det - 1
    out = dfull[c:].copy()
    out[1:] += dfull[:c][::-1]
    return out


def learned_fbp(prm, sino, n, angles=None, C=16):
    """sino (B, nang, ndet) -> r0, the learned-filter FBP."""
    g = unpack(prm, C)
    h = full_taps(g["taps"])
    sino = np.asarray(sino, np.float64)
    r0 = np.stack([std_fbp.backproject(std_fbp.filter_rows(s, h), n, angles) for s in sino])
    return r0


def loss_and_grad(prm, sino, target, bf=True, angles=None, C=16):
    """Full model: loss = mean((out - target)^2) and its gradient w.r.t. every
    parameter (taps included)."""
    sino = np.asarray(sino, np.float64)
    target = np.asarray(target, np.float64)
    B, n_angles, n_det = sino.shape
    n = target.shape[-1]
    r0 = learned_fbp(prm, sino, n, angles, C)
    cache = cnn_forward(prm, r0, bf, C)
    diff = cache["out"] - target
    count = diff.size
    loss = float((diff ** 2).sum() / count)
    dout = 2.0 * diff / count
    grads, dr0 = cnn_backward(prm, cache, dout, bf, C)
    dfull = np.zeros(2 * n_det - 1)
    for b in range(B):
        dq = backproject_adjoint(dr0[b], n_angles, n_det, angles)
        dfull += filter_grad(dq, sino[b])
    grads["taps"] = fold_taps(dfull)
    return loss, pack(grads, n_det, C), {"cache": cache, "dr0": dr0, "dout": dout}


def psnr(rec, target):
    """10 log10(range^2 / mse), range = max - min of the target, per image."""
    rec = np.asarray(rec, np.float64)
    target = np.asarray(target, np.float64)
    if target.ndim == 2:
        rec, target = rec[None], target[None]
    out = []
    for r, t in zip(rec, target):
        rng = t.max() - t.min()
        out.append(10.0 * np.log10(rng * rng / np.mean((r - t) ** 2)))
    return np.array(out)
