"""Standard-geometry projector / FBP on the GPU against the FP64 CPU oracle
(oracle/std_fbp.py): bar 1e-4 relative to the largest magnitude."""
import numpy as np
import pytest

from oracle import std_fbp
from skewengine_amd import synth
from skewengine_amd.fbp import FBP

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("n,nang,ndet", [(64, 30, 91), (128, 45, 183), (96, 60, 137)])
def test_project_filter_backproject_match_fp64(n, nang, ndet):
    img = synth.ellipse_phantom(n, seed=n)
    with FBP(n, nang, ndet, max_batch=2) as f:
        sino = f.radon(np.stack([img, img[::-1].copy()]))
        ref = std_fbp.project(img, nang, ndet)
        assert _rel(sino[0], ref) <= 1e-4
        assert _rel(sino[1], std_fbp.project(img[::-1], nang, ndet)) <= 1e-4
        rec = f.fbp(sino[0])
        ref_rec = std_fbp.fbp(sino[0].astype(np.float64), n)
        assert _rel(rec, ref_rec) <= 1e-4
        # filtered rows alone
        f.upload(1, sino[:1])
        f.filter(1)
        q = f.download(2, 1)[0]
        assert _rel(q, std_fbp.filter_rows(sino[0].astype(np.float64), std_fbp.ramlak_taps(ndet, n))) <= 1e-4


def test_c2_geometry_subset_of_angles():
    # 512^2, 180 angles, 725 bins: compare a subset of projection angles in FP64
    img = synth.ellipse_phantom(512, seed=2)
    with FBP(512, 180, 725) as f:
        sino = f.radon(img)
        idx = [0, 1, 44, 45, 90, 133, 179]
        ref = std_fbp.project(img, 180, 725, angle_idx=idx)
        assert _rel(sino[idx], ref) <= 1e-4
        rec = f.fbp(sino)
    ref_rec = std_fbp.fbp(sino.astype(np.float64), 512)
    assert _rel(rec, ref_rec) <= 1e-4


def test_learned_taps_path_and_reset():
    n, nang, ndet = 64, 30, 91
    img = synth.ellipse_phantom(n, seed=7)
    with FBP(n, nang, ndet) as f:
        sino = f.radon(img)
        taps = np.random.default_rng(0).standard_normal(2 * ndet - 1).astype(np.float32)
        f.set_taps(taps)
        assert np.array_equal(f.taps(), taps)
        rec = f.fbp(sino)
        assert _rel(rec, std_fbp.fbp(sino.astype(np.float64), n, taps=taps.astype(np.float64))) <= 1e-4
        f.set_taps(None)
        assert np.allclose(f.taps(), std_fbp.ramlak_taps(ndet, n), rtol=1e-6)
