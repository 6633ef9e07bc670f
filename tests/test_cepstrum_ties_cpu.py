"""The cepstrum tie checker (tests/cepstrum_ties.py) on the CPU: it accepts
a flat-spectrum window whose peak-rule samples are rounding noise, and it
rejects a planted wrong value on a cell whose peak rule is well separated."""
import numpy as np
import pytest

import cepstrum_ties as ct
import oracle
from skewengine_amd import synth


def test_flat_spectrum_window_is_a_tie():
    # one nonzero sample in the window: |X_k| is constant, ln|X|^2 is flat and
    # every C_k with k > 0 is rounding noise, so 0 and the oracle's tiny value
    # are both correct results
    row = np.zeros(40, np.float32)
    row[5] = 0.37
    v, logps, ceps = oracle.line_cepstrum_ex(6, row, 5, 40)
    assert np.ptp(logps) <= 1e-12  # flat up to rounding
    delta = ct.error_bound(ct.window(row, 6, 5, 40), logps)
    assert np.all(np.abs(ceps[1:]) <= delta)
    assert ct.is_tie(ceps, delta, 0.0)
    assert ct.is_tie(ceps, delta, v)


@pytest.fixture(scope="module")
def sector46():
    img = synth.blurred_texture(96, 0.6, seed=4)[0]
    return oracle.sector(oracle.CEPSTRUM, img, 46)


def test_oracle_against_itself_has_no_mismatch(sector46):
    p = sector46["p"]
    SO = sector46["SO"][:p.skewHeight * p.dim_i].reshape(p.skewHeight, p.dim_i).copy()
    bad, cells = ct.check_sector(SO, sector46)
    assert bad == 0 and cells > 0


def test_planted_wrong_value_is_rejected(sector46):
    p = sector46["p"]
    H, di = p.skewHeight, p.dim_i
    SO = sector46["SO"][:H * di].reshape(H, di).copy()
    S = sector46["S"][:H * di].reshape(H, di)
    # a cell with a clearly nonzero value whose peak rule is not fragile
    for i, j in zip(*np.nonzero(SO > 1.0)):
        f, l = int(sector46["first"][i]), int(sector46["last"][i])
        _, logps, ceps = oracle.line_cepstrum_ex(int(j), S[i], f, l)
        if not ct.peak_rule_fragile(ceps, ct.error_bound(ct.window(S[i], int(j), f, l), logps)):
            break
    else:
        pytest.fail("no well-separated cell found")
    SO[i, j] *= 1.5
    with pytest.raises(AssertionError):
        ct.check_sector(SO, sector46)
    SO[i, j] = 0.0  # zero is a legal peak-rule result only at a tie
    with pytest.raises(AssertionError):
        ct.check_sector(SO, sector46)


def test_peak_rule_replay_sees_the_decisive_comparison():
    # samples 5 > 4 is the first ascent: with v4, v5 apart by 1e-3 the rule
    # is certain at delta 1e-6 and uncertain at delta 1e-3
    ceps = np.zeros(16)
    ceps[4], ceps[5], ceps[9] = -0.100, -0.101, -0.05
    assert not ct.peak_rule_fragile(ceps, 1e-6)
    assert ct.peak_rule_fragile(ceps, 1e-3)
    # a zero sample from Re C_k near +0 makes the comparison against it uncertain
    ceps = np.full(16, 0.5)
    ceps[6] = 1e-9
    assert ct.peak_rule_fragile(ceps, 1e-8)
    assert not ct.peak_rule_fragile(ceps, 1e-10)
