"""Sample points of the local Radon transform (saveSampleData,
src/global.cpp:47-56; recorded by pointRadon, src/kernelRADN.h:164-170;
written by showResultsRADN, src/kernelRADN.h:263-382) on the GPU against
the oracle's restatement of the same four vectors."""
import os
import tempfile

import numpy as np
import pytest

import oracle
import skewengine_amd as ske
from skewengine_amd import _lib, synth

pytestmark = pytest.mark.gpu

N = 96
POINTS = [(20, 24), (5, 40), (60, 70), (200, 200)]  # the last lies outside the image


@pytest.fixture(scope="module")
def phantom():
    return synth.ellipse_phantom(N, seed=13)


@pytest.fixture(scope="module")
def record(phantom):
    with ske.SkewEngine(phantom) as e:
        e.set_option(_lib.OPT_RADON_LOCAL, 1)
        e.set_sample_points(POINTS)
        e.run(ske.KERNEL_RADON)
        return e.sample_data()


def test_record_matches_oracle_vectors(phantom, record):
    assert record.shape == (180, len(POINTS), 4, 64)
    assert not np.any(record[:, 3])  # outside the image: never visited
    checked = 0
    for a in list(range(0, 180, 7)) + [45, 46, 90, 135]:
        sec = oracle.sector(oracle.IDENTITYIMG, phantom, a)  # same skewed input, maps and extents
        p = sec["p"]
        for k, (x, y) in enumerate(POINTS[:3]):
            i, j = oracle.skewed_cell(p, sec["target"], x, y, N, N)
            inside = 0 <= i < p.skewHeight and sec["first"][i] <= j < sec["last"][i]
            if not inside:
                assert not np.any(record[a, k]), (a, x, y)
                continue
            ref = oracle.point_radon_samples(sec, i, j)
            scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1e-30)
            assert np.all(np.abs(record[a, k] - ref) <= 1e-6 * scale), (a, x, y)
            checked += 1
    assert checked > 60


def test_per_sector_path_records_the_same(phantom, record):
    with ske.SkewEngine(phantom) as e:
        e.set_option(_lib.OPT_RADON_LOCAL, 1)
        e.set_sample_points(POINTS)
        e.set_kernel(ske.KERNEL_RADON, 0.15)
        sectors = (3, 45, 100, 170)
        for a in sectors:
            e.skew(a)
            e.kernel()
        per = e.sample_data()
    for a in range(180):
        if a in sectors:
            assert np.array_equal(per[a], record[a]), a
        else:
            assert not np.any(per[a]), a


def test_run_writes_the_reference_dump(phantom, record):
    with tempfile.TemporaryDirectory() as d:
        fin = os.path.join(d, "img.flt")
        phantom.tofile(fin)
        pts = ";".join("%d,%d" % p for p in POINTS[:3])
        rc = ske.run_cli(["--kernel", "radon", "--local", "--input", fin, "--sample-dir", d, "--samples", pts])
        assert rc == 0
        for k in range(3):
            for v, kind in enumerate("oilc"):
                got = np.loadtxt(os.path.join(d, "punto%d%s.txt" % (k + 1, kind)))
                assert got.shape == (360, 32)
                vec = record[:, k, v]
                second = vec[:, 32 - np.arange(32)]
                first = second if kind == "l" else vec[:, 32 + np.arange(32)]
                exp = np.concatenate([first, second])
                assert np.allclose(got, exp, rtol=1e-5, atol=1e-30), (k, kind)
