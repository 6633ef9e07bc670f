"""Pins of the oracle's max-mode deskew and reduce with hand-worked values.

deskew(isMax=1) (src/skewEngine.cuh:775-783): for every output cell
    O = max(O, max(SO[i+t_j, j], SO[i+t_j+1, j]))     (std::max)
    O = a + 1000 * (int)(O / 1000)                    (float quotient, C truncation)
reduce(isMax=1) (src/skewEngine.cuh:885-895): out = MAX(out, o_k[map_k(i, j)])
for k = 0..3 in that order with MAX(a, b) = a > b ? a : b.

Every expected number below was worked out by hand from those two rules
(the working is in the comments), not by re-running a formula.
"""
import numpy as np

import oracle


def _params(a, dim_i, dim_o):
    p = oracle.Params()
    p.a, p.dim_i, p.dim_o = a, dim_i, dim_o
    return p


# skewed output of a 3-column sector: rows 0..3, target = [0, 1, 1], weight unused in max mode
SO = np.array([[12345.6, 100.0, 7.0],
               [-2000.0, 2999.9, -50.0],
               [-1800.0, 1000.0, 999.99],
               [0.0, 4242.0, 0.0]], np.float32)
TARGET = np.array([0, 1, 1], np.int32)
WEIGHT = np.array([0.0, 0.25, 0.75], np.float32)


def test_deskew_max_encode_hand_values():
    O = np.array([0.0, 3000.0, 61000.0, -1500.0, 0.0, 0.0], np.float32)
    oracle.deskew(_params(37, 3, 2), TARGET, WEIGHT, SO.reshape(-1), O, is_max=1)
    # (0,0) t=0: max(0, 12345.6, -2000) = 12345.6 -> 12.3456 -> 12 -> 37 + 12000
    # (0,1) t=1: max(3000, 2999.9, 1000) = 3000 -> 3 -> 37 + 3000
    # (0,2) t=1: max(61000, -50, 999.99) = 61000 -> 61 -> 37 + 61000
    # (1,0) t=0: max(-1500, -2000, -1800) = -1500 -> -1.5 -> -1 (toward zero) -> 37 - 1000
    # (1,1) t=1: max(0, 1000, 4242) = 4242 -> 4 -> 37 + 4000
    # (1,2) t=1: max(0, 999.99, 0) = 999.99 -> 0.99999 -> 0 -> 37
    assert O.tolist() == [12037.0, 3037.0, 61037.0, -963.0, 4037.0, 37.0]
    # a later sector re-encodes: the angle digits are replaced, the thousands kept
    oracle.deskew(_params(38, 3, 2), TARGET, WEIGHT, SO.reshape(-1), O, is_max=1)
    # (0,0): max(12037, 12345.6) = 12345.6 -> 12038; (0,1): max(3037, 2999.9) -> 3038
    # (1,0): max(-963, -1800) = -963 -> -0.963 -> 0 -> 38; (1,2): max(37, 999.99) = 999.99 -> 38
    assert O.tolist() == [12038.0, 3038.0, 61038.0, 38.0, 4038.0, 38.0]


def test_deskew_sum_mode_hand_values():
    O = np.zeros(6, np.float32)
    oracle.deskew(_params(5, 3, 2), TARGET, WEIGHT, SO.reshape(-1), O, is_max=0)
    # (0,0): 1*12345.6 + 0*(-2000); (0,1): 0.75*2999.9 + 0.25*1000 = 2249.925 + 250
    # (1,2): 0.25*999.99 + 0.75*0 = 249.9975 (each product rounded to float first)
    exp01 = np.float32(np.float32(0.75) * np.float32(2999.9)) + np.float32(np.float32(0.25) * np.float32(1000.0))
    assert O[0] == np.float32(12345.6)
    assert O[1] == exp01 and abs(float(O[1]) - 2499.925) < 1e-3
    assert abs(float(O[5]) - 249.9975) < 1e-4


def test_reduce_max_hand_values():
    # 2x2 grid, ci = 1-i, cj = 1-j.  Reads for output cell (i, j):
    #   o0[2i+j], o1[2cj+ci], o2[2j+ci], o3[2i+cj]
    #   (0,0): o0[0] o1[3] o2[1] o3[1]     (0,1): o0[1] o1[1] o2[3] o3[0]
    #   (1,0): o0[2] o1[2] o2[0] o3[3]     (1,1): o0[3] o1[0] o2[2] o3[2]
    o0 = np.array([50, 1, 1, 1], np.float32)
    o1 = np.array([2, 60, 2, 2], np.float32)
    o2 = np.array([70, 3, 3, 3], np.float32)
    o3 = np.array([4, 4, 80, 4], np.float32)
    out = np.array([-1, -1, -1, -1], np.float32)
    oracle.reduce(2, 2, out, [o0, o1, o2, o3], is_max=1)
    assert out.tolist() == [50.0, 60.0, 70.0, 80.0]
    # absent orientations are skipped and a larger starting value survives
    out = np.array([-1, 100, -1, -1], np.float32)
    oracle.reduce(2, 2, out, [None, o1, None, o3], is_max=1)
    # (0,0): o1[3]=2, o3[1]=4 -> 4; (0,1): 100; (1,0): o1[2]=2, o3[3]=4 -> 4; (1,1): o1[0]=2, o3[2]=80
    assert out.tolist() == [4.0, 100.0, 4.0, 80.0]


def test_reduce_max_order_shows_in_signed_zeros():
    # MAX(a, b) = a > b ? a : b keeps b on ties, so with out = +0, o0 = -0 the
    # result is -0; a later +0 from o1 turns it back into +0 (o0 before o1).
    z, nz = np.float32(0.0), np.float32(-0.0)
    out = np.array([z], np.float32)
    oracle.reduce(1, 1, out, [np.array([nz], np.float32), None, None, None], is_max=1)
    assert np.signbit(out[0])
    out = np.array([z], np.float32)
    oracle.reduce(1, 1, out, [np.array([nz], np.float32), np.array([z], np.float32), None, None], is_max=1)
    assert not np.signbit(out[0])
    out = np.array([z], np.float32)
    oracle.reduce(1, 1, out, [np.array([z], np.float32), np.array([nz], np.float32), None, None], is_max=1)
    assert np.signbit(out[0])


def test_deskew_max_keeps_last_angle_on_equal_thousands():
    # two sectors whose maxima fall in the same thousand: the later angle wins
    SO1 = np.array([[5100.0], [0.0]], np.float32).reshape(-1)
    SO2 = np.array([[5900.0], [0.0]], np.float32).reshape(-1)
    t = np.zeros(1, np.int32)
    w = np.zeros(1, np.float32)
    O = np.zeros(1, np.float32)
    oracle.deskew(_params(10, 1, 1), t, w, SO1, O, is_max=1)  # 5100 -> 5010
    assert O[0] == 5010.0
    oracle.deskew(_params(99, 1, 1), t, w, SO2, O, is_max=1)  # max(5010, 5900) = 5900 -> 5099
    assert O[0] == 5099.0
    oracle.deskew(_params(120, 1, 1), t, w, SO1, O, is_max=1)  # max(5099, 5100) = 5100 -> 5120
    assert O[0] == 5120.0
