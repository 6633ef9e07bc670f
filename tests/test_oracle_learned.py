"""Pins of the FP64 learned-FBP oracle (oracle/learned_ref.py): central
finite differences of the loss for every parameter block and the taps, the
back-projector adjoint identity, the filter gradient, and bf16 rounding."""
import numpy as np
import pytest

from oracle import learned_ref as L
from oracle import std_fbp


def _problem(n=32, na=6, nd=45, seed=0, scale=0.2):
    rng = np.random.default_rng(seed)
    prm = np.zeros(L.TAPS + nd)
    prm[:L.TAPS] = rng.standard_normal(L.TAPS) * scale
    prm[L.TAPS:] = std_fbp.ramlak_taps(nd, n)[nd - 1:]
    sino = rng.standard_normal((2, na, nd))
    tgt = rng.standard_normal((2, n, n))
    return prm, sino, tgt


def test_gradient_matches_central_differences():
    prm, sino, tgt = _problem()
    _, g, _ = L.loss_and_grad(prm, sino, tgt, bf=False)
    rng = np.random.default_rng(1)
    idx = []
    for name, (o, s) in L.OFF.items():
        idx += list(o + rng.choice(s, size=min(s, 3), replace=False))
    idx += [L.TAPS, L.TAPS + 1, L.TAPS + 3, L.TAPS + 40]
    e = 1e-6
    for i in idx:
        p1, p2 = prm.copy(), prm.copy()
        p1[i] += e
        p2[i] -= e
        fd = (L.loss_and_grad(p1, sino, tgt, bf=False)[0] - L.loss_and_grad(p2, sino, tgt, bf=False)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_dr0_matches_central_differences():
    prm, _, _ = _problem()
    rng = np.random.default_rng(2)
    r0 = rng.standard_normal((1, 16, 16))
    tgt = rng.standard_normal((1, 16, 16))

    def loss(r):
        out = L.cnn_forward(prm, r, bf=False)["out"]
        return ((out - tgt) ** 2).mean()

    cache = L.cnn_forward(prm, r0, bf=False)
    dout = 2 * (cache["out"] - tgt) / tgt.size
    _, dr0 = L.cnn_backward(prm, cache, dout, bf=False)
    for (i, j) in [(0, 0), (3, 7), (15, 15), (8, 0)]:
        e = 1e-6
        rp, rm = r0.copy(), r0.copy()
        rp[0, i, j] += e
        rm[0, i, j] -= e
        fd = (loss(rp) - loss(rm)) / (2 * e)
        assert abs(fd - dr0[0, i, j]) <= 1e-6 * max(1.0, abs(fd))


def test_backproject_adjoint_identity():
    rng = np.random.default_rng(3)
    n, na, nd = 40, 7, 61
    q = rng.standard_normal((na, nd))
    r = rng.standard_normal((n, n))
    lhs = np.sum(std_fbp.backproject(q, n) * r)
    rhs = np.sum(q * L.backproject_adjoint(r, na, nd))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)


def test_filter_grad_is_the_adjoint_of_the_taps():
    rng = np.random.default_rng(4)
    nd = 23
    p = rng.standard_normal((3, nd))
    dq = rng.standard_normal((3, nd))
    h = rng.standard_normal(2 * nd - 1)
    # <dq, filter(p, h)> is linear in h: its gradient is filter_grad(dq, p)
    dh = L.filter_grad(dq, p)
    for m in (0, 5, nd - 1, 2 * nd - 2):
        e = np.zeros_like(h)
        e[m] = 1.0
        val = np.sum(dq * (std_fbp.filter_rows(p, h + e) - std_fbp.filter_rows(p, h)))
        assert abs(val - dh[m]) <= 1e-10 * max(1.0, abs(val))
    # symmetric folding: d/dsym[k] of sum_m dh[m] sym[|m|]
    sym = L.fold_taps(dh)
    assert np.isclose(sym[0], dh[nd - 1]) and np.isclose(sym[3], dh[nd - 1 + 3] + dh[nd - 1 - 3])


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -9, 3.0e38, 1e-40])
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, -2.5, 1.0, np.float32(3.0e38), 1e-40])
    got = L.bf16(vals)
    assert not np.any(got.astype(np.float32).view(np.uint32) & 0xFFFF)  # representable in bf16
    assert got[0] == 1.0 and got[1] == 1.0  # tie to even
    assert got[2] == 1.0 + 2 ** -6          # tie to even (odd lower neighbour rounds up)
    assert got[3] == -2.5 and got[4] == 1.0
    assert abs(got[5] - want[5]) / want[5] < 2 ** -8
    assert abs(got[6] - 1e-40) <= 2 ** -133  # half a bf16 ulp among fp32 denormals


@pytest.mark.parametrize("bf", [False, True])
def test_zero_output_layer_gives_plain_fbp(bf):
    n, na, nd = 32, 8, 45
    rng = np.random.default_rng(5)
    prm = np.zeros(L.TAPS + nd)
    prm[:L.OFF["W4"][0]] = rng.standard_normal(L.OFF["W4"][0]) * 0.3
    prm[L.TAPS:] = std_fbp.ramlak_taps(nd, n)[nd - 1:]
    sino = rng.standard_normal((1, na, nd))
    r0 = L.learned_fbp(prm, sino, n)
    out = L.cnn_forward(prm, r0, bf)["out"]
    assert np.allclose(out[0], std_fbp.fbp(sino[0], n), rtol=0, atol=1e-9 * np.abs(out).max())


def test_gradient_matches_central_differences_32_channels():
    """The same pin for the 32-channel layout (W1[32][9] ... W4[9][32] b4 taps)."""
    off, taps = L.layout(32)
    assert taps == 19105 and off["W2"] == (320, 9216)
    n, na, nd = 32, 6, 45
    rng = np.random.default_rng(5)
    prm = np.zeros(taps + nd)
    prm[:taps] = rng.standard_normal(taps) * 0.1
    prm[taps:] = std_fbp.ramlak_taps(nd, n)[nd - 1:]
    sino = rng.standard_normal((1, na, nd))
    tgt = rng.standard_normal((1, n, n))
    _, g, _ = L.loss_and_grad(prm, sino, tgt, bf=False, C=32)
    assert g.size == taps + nd
    idx = []
    for name, (o, s) in off.items():
        idx += list(o + rng.choice(s, size=min(s, 2), replace=False))
    idx += [taps, taps + 7]
    e = 1e-6
    for i in idx:
        p1, p2 = prm.copy(), prm.copy()
        p1[i] += e
        p2[i] -= e
        fd = (L.loss_and_grad(p1, sino, tgt, bf=False, C=32)[0] - L.loss_and_grad(p2, sino, tgt, bf=False, C=32)[0]) \
            / (2 * e)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_layout_sizes():
    for C, taps in ((16, 4945), (32, 19105)):
        off, t = L.layout(C)
        assert t == taps
        assert off["W1"] == (0, 9 * C) and off["b1"] == (9 * C, C)
        assert off["W2"] == (10 * C, 9 * C * C) and off["W4"][1] == 9 * C and off["b4"] == (t - 1, 1)
