"""Pins of the oracle's index maps (computeParams / skewShape / prepare /
reduce, src/skewEngine.cuh:461-596,861-987) that do not re-type its code:
hand-computed values, invariants, an independent closed form of the
`last[]` parity rule, and the orientation maps as numpy flips/transposes."""
import math

import numpy as np
import pytest

import oracle

SIZES = [(500, 500), (64, 64), (7, 7), (512, 512)]


def test_fangle_square_is_45():
    # atan2(n,n)*180/pi rounds to exactly 45.0f for a square grid.
    for n in (5, 64, 500, 4000):
        assert oracle.fangle(n, n) == 45.0


def test_sector_types_square():
    # 0-45 type 0, 46-90 type 1, 91-134 type 2, 135-179 type 3 (46/45/44/45 sectors).
    types = [oracle.compute_params(100, 100, a)[0].sectorType for a in range(180)]
    assert types == [0] * 46 + [1] * 45 + [2] * 44 + [3] * 45
    for a in range(180):
        p = oracle.compute_params(100, 100, a)[0]
        expect = {0: a, 1: 90 - a, 2: a - 90, 3: 180 - a}[p.sectorType]
        assert p.newAngle == expect


def test_hand_values_45_degrees():
    # tan(45 deg) in double is 0.9999999999999999 < 1: t_j = j-1 for j >= 1 and
    # the weight rounds to 1.0f.
    p, t, w, first, last = oracle.compute_params(500, 500, 45)
    assert p.newAngle == 45 and p.sectorType == 0
    assert p.skewness < 1.0
    assert list(t[:6]) == [0, 0, 1, 2, 3, 4]
    assert w[0] == 0.0 and np.all(w[1:500] == np.float32(1.0))
    assert p.offset == 498 and p.skewHeight == 999


def test_hand_values_30_degrees():
    p, t, w, first, last = oracle.compute_params(500, 500, 30)
    # tan(30 deg) = 1/sqrt(3) = 0.57735...; 499 * 0.57735 = 288.09...
    assert abs(p.skewness - 1 / math.sqrt(3)) < 1e-15
    assert p.offset == 288 and p.skewHeight == 500 + 288 + 1
    assert abs(p.scale - 2 / math.sqrt(3)) < 1e-15
    assert t[10] == 5 and abs(w[10] - np.float32(10 / math.sqrt(3) - 5)) < 1e-6


def test_last_parity_quirk_example():
    # 500^2 at 10 deg: row 2 is one cell short (last 17 instead of 18).
    p, t, w, first, last = oracle.compute_params(500, 500, 10)
    assert last[2] == 17
    true_last = max(j for j in range(500) if t[j] in (1, 2)) + 1
    assert true_last == 18


def closed_form_extents(p, t):
    """first/last from the explicit rule (no sequential update loop).

    Rows r < offset: the columns j with t_j in {r-1, r} form [alpha, J]; the
    reference's `if (j > last) last = j+1` fires at alpha, alpha+2, ... when
    alpha > 0 and at 1, 3, ... when alpha == 0.  Rows >= offset keep dim_l.
    first is the smallest j with t_j + dim_o - 1 in {r-1, r} for r >= dim_o,
    else 0."""
    H, dim_o, dim_l, off = p.skewHeight, p.dim_o, p.dim_l, p.offset
    t = t[:p.dim_i].astype(np.int64)
    first = np.zeros(H, np.int64)
    last = np.zeros(H, np.int64)
    for r in range(H):
        if r >= off:
            last[r] = dim_l
        else:
            js = np.nonzero((t == r) | (t == r - 1))[0]
            if len(js) == 0:
                last[r] = 0
            else:
                a, J = js[0], js[-1]
                assert np.all(np.diff(js) == 1)
                if a > 0:
                    last[r] = a + 2 * ((J - a) // 2) + 1
                else:
                    last[r] = 2 * ((J - 1) // 2) + 2 if J >= 1 else 0
        if r < dim_o:
            first[r] = 0
        else:
            js = np.nonzero((t + dim_o - 1 == r) | (t + dim_o - 1 == r - 1))[0]
            first[r] = js[0] if len(js) else dim_l
    return first, last


@pytest.mark.parametrize("dims", SIZES)
def test_extents_closed_form_all_sectors(dims):
    dimx, dimy = dims
    for a in range(180):
        p, t, w, first, last = oracle.compute_params(dimx, dimy, a)
        f2, l2 = closed_form_extents(p, t)
        H = p.skewHeight
        assert np.array_equal(first[:H].astype(np.int64), f2), a
        assert np.array_equal(last[:H].astype(np.int64), l2), a


@pytest.mark.parametrize("dims", SIZES)
def test_map_invariants(dims):
    dimx, dimy = dims
    for a in range(180):
        p, t, w, first, last = oracle.compute_params(dimx, dimy, a)
        ti = t[:p.dim_i]
        wi = w[:p.dim_i]
        assert ti[0] == 0
        assert np.all(np.diff(ti) >= 0) and np.all(np.diff(ti) <= 1)
        assert np.all(wi >= 0) and np.all(wi <= 1)
        assert p.offset == ti[-1]
        assert p.skewHeight == (p.dim_o + p.offset + 1 if p.newAngle else p.dim_o)
        # every cell in [first, last) has at least one source tap; first is
        # the true first column, last at most one short of the true extent
        H = p.skewHeight
        r = np.arange(H)[:, None]
        has_tap = (r >= ti[None, :]) & (r <= ti[None, :] + p.dim_o)
        j = np.arange(p.dim_i)[None, :]
        inside = (j >= first[:H, None].astype(int)) & (j < last[:H, None].astype(int))
        assert not np.any(inside & ~has_tap)
        any_tap = has_tap.any(axis=1)
        true_first = np.argmax(has_tap, axis=1)
        true_last = p.dim_i - np.argmax(has_tap[:, ::-1], axis=1)
        assert np.all(first[:H][any_tap] == true_first[any_tap])
        assert np.all(last[:H][any_tap].astype(int) >= true_last[any_tap] - 1)


def test_prepare_is_flip_transpose():
    rng = np.random.default_rng(1)
    img = rng.standard_normal((37, 53)).astype(np.float32)  # dimy=37, dimx=53
    in0, in1, in2, in3 = oracle.prepare(img)
    assert np.array_equal(in1.reshape(53, 37), img[::-1, ::-1].T)
    assert np.array_equal(in2.reshape(53, 37), img[::-1, :].T)
    assert np.array_equal(in3.reshape(37, 53), img[:, ::-1])


def test_reduce_inverts_prepare():
    # reduce() maps each orientation buffer back: with one buffer present at a
    # time, reduce(prepare(x)) reproduces x exactly.
    rng = np.random.default_rng(2)
    img = rng.standard_normal((37, 53)).astype(np.float32)
    bufs = oracle.prepare(img)
    for k in range(4):
        Os = [None] * 4
        Os[k] = np.ascontiguousarray(bufs[k])
        out = np.zeros(37 * 53, np.float32)
        oracle.reduce(53, 37, out, Os)
        assert np.array_equal(out.reshape(37, 53), img), k


def test_work_counts_match_independent_restatement():
    # Skewed-cell and ray-step counts computed by an independent restatement
    # of the maps (written separately from the oracle).
    assert oracle.work(500, 500) == pytest.approx((45068242, 1.9179e10), rel=1e-4)
    c, s = oracle.work(512, 512)
    assert c == 47255794
    c, s = oracle.work(1000, 1000)
    assert c == 180136434
