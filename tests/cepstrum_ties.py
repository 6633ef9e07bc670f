"""Checker of cepstrum mismatches (test helper; not collected).

The cepstral blur operator (src/kernelBLUR.h:153-228) ends in a peak rule
(getBlurParameter2, :88-110) that picks one of 16 samples
v_k = -min(0, Re C_k) by comparisons.  Where a comparison the rule makes is
closer than the rounding error of the FP64 computation, two correct FP64
implementations may pick different samples; elsewhere they must agree to
1e-4.  `check_sector` compares one sector's skewOutput with the oracle and
proves every cell outside rtol 1e-4 to be such a tie:

* the error bound.  X = DFT64(window)/8 has |X_k| <= E = sum|x|/8 and each
  X_k carries an absolute error below C*eps*E (C = 64: a generous allowance
  for the six radix-2 stages of the reference's FFT and the rotations of a
  sliding DFT); ln|X_k|^2 then errs by at most 2*C*eps*E/|X_k| plus its own
  rounding, and C_k = DFT32(ln|X|^2)/sqrt(32) adds at most 32*eps*sum|ln|X|^2|
  before the scaling.  Delta, the sum, bounds the error of every Re C_k, so
  each v_k lies in [max(0, -(C_k + Delta)), max(0, -(C_k - Delta))].
* a tie: replaying the peak rule on the oracle's values, some comparison it
  makes has overlapping intervals (it could go either way), and the GPU value
  is one of the samples (within Delta) or zero.
"""
import numpy as np

import oracle

EPS = 2.0 ** -52
C_DFT = 64.0


def error_bound(window, logps):
    """Bound of the rounding error of the 16 values Re C_k of one cell."""
    E = np.sum(np.abs(window.astype(np.float64))) / 8.0
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        absX = np.exp(logps / 2.0)
        dlog = 2.0 * C_DFT * EPS * E / absX + 4.0 * EPS * np.abs(logps)
        d = (np.sum(dlog) + 32.0 * EPS * np.sum(np.abs(logps))) / np.sqrt(32.0)
    return float(d) if np.isfinite(d) else np.inf


def peak_rule_fragile(ceps, delta):
    """True when a comparison getBlurParameter2 makes on the samples
    -min(0, ceps) can flip within +-delta of every Re C_k."""
    v = -np.minimum(0.0, ceps)
    lo = np.maximum(0.0, -(ceps + delta))
    hi = np.maximum(0.0, -(ceps - delta))

    def uncertain(a, b):  # the comparison v[a] > v[b]
        return not (lo[a] > hi[b] or hi[a] <= lo[b])

    i = 5
    while i < 16:  # the first ascent at or after 5
        if uncertain(i, i - 1):
            return True
        if v[i] > v[i - 1]:
            break
        i += 1
    if i == 16:
        return False
    cur = i - 1  # val = v[i-1]; v[i] itself is skipped
    for m in range(i + 1, 16):
        if uncertain(m, cur):
            return True
        if v[m] > v[cur]:
            cur = m
    return False


def is_tie(ceps, delta, gpu_value):
    samples = -np.minimum(0.0, ceps)
    g = float(gpu_value) / 10000.0
    tol = delta + 1e-6 * abs(g)  # (float)(V * 10000) rounds to 24 bits
    candidate = abs(g) <= tol or np.min(np.abs(samples[4:] - g)) <= tol
    return bool(candidate and peak_rule_fragile(ceps, delta))


def window(row, j, first, last):
    w = np.zeros(64, np.float32)
    for m in range(64):
        idx = j - 32 + m
        if first <= idx < last:
            w[m] = row[idx]
    return w


def check_sector(SO, ref):
    """SO: the GPU skewOutput rows (H x dim_i) of one sector; ref: the
    oracle.sector() result for it.  Asserts that every cell outside rtol
    1e-4 is a proven tie; returns (mismatches, valid cells)."""
    p = ref["p"]
    H, di = p.skewHeight, p.dim_i
    S = ref["S"][:H * di].reshape(H, di)
    rSO = ref["SO"][:H * di].reshape(H, di)
    bad = ~np.isclose(SO, rSO, rtol=1e-4, atol=0.0)
    for i, j in zip(*np.nonzero(bad)):
        f, l = int(ref["first"][i]), int(ref["last"][i])
        assert f <= j < l, ("mismatch outside the valid cells", i, j)
        v, logps, ceps = oracle.line_cepstrum_ex(int(j), S[i], f, l)
        assert np.float32(v) == rSO[i, j]
        delta = error_bound(window(S[i], int(j), f, l), logps)
        assert is_tie(ceps, delta, SO[i, j]), (p.a, int(i), int(j), float(SO[i, j]), float(rSO[i, j]),
                                               ceps.tolist(), delta)
    cells = int(sum(max(int(ref["last"][i]) - int(ref["first"][i]), 0) for i in range(H)))
    return int(bad.sum()), cells


def deskew_reduce(n, sectors, is_max=1):
    """The oracle's deskew (max-encode) and reduce of per-sector skewOutputs
    [(params, target, weight, SO)], in the order given."""
    accs = [None, None, None, None]
    for p, target, weight, SO in sectors:
        t = p.sectorType
        if accs[t] is None:
            accs[t] = np.zeros(n * n, np.float32)
        full = np.zeros(2 * n * n + n, np.float32)
        full[:SO.size] = SO.reshape(-1)
        oracle.deskew(p, target, weight, full, accs[t], is_max)
    out = np.zeros(n * n, np.float32)
    oracle.reduce(n, n, out, accs, is_max)
    return out.reshape(n, n)


def gpu_sectors_checked(img, sectors=range(180)):
    """Run each sector through the GPU per-sector path, prove every
    mismatch with the oracle a tie, and return the oracle-side tables with
    the GPU skewOutputs [(params, target, weight, SO)] for deskew_reduce."""
    import skewengine_amd as ske
    n = img.shape[0]
    out = []
    with ske.SkewEngine(img) as e:
        e.set_kernel(ske.KERNEL_CEPSTRUM, 0.15)
        for a in sectors:
            e.skew(a)
            e.kernel()
            _, SO = e.skew_buffers()
            ref = oracle.sector(oracle.CEPSTRUM, img, a)
            check_sector(SO, ref)
            p, target, weight, _, _ = oracle.compute_params(n, n, a)
            out.append((p, target, weight, SO.copy()))
    return out
