"""run() on a PNG input (configureRADN, src/kernelRADN.h:78-88: the grey
level is R+G+B of the 8-bit RGBA decode, src/helper.h:167-186) for the
Radon and identity-image operators, against the oracle on that grey image;
the result images salida1.png / salida2.png are written (src/main.h:124-141)."""
import os

import numpy as np
import pytest

import oracle
import skewengine_amd as ske
from skewengine_amd import synth

pytestmark = pytest.mark.gpu

Image = pytest.importorskip("PIL.Image")


@pytest.mark.parametrize("kernel,oracle_id", [("radon", oracle.RADON), ("identityimg", oracle.IDENTITYIMG)])
def test_png_input_through_run(tmp_path, kernel, oracle_id):
    n = 72
    base = synth.ellipse_phantom(n, seed=31)
    base = (base - base.min()) / (base.max() - base.min())
    rng = np.random.default_rng(4)
    rgb = np.stack([base * 255, base[::-1] * 200, rng.random((n, n)) * 90], axis=2).astype(np.uint8)
    png = str(tmp_path / "in.png")
    Image.fromarray(rgb, "RGB").save(png)
    grey = rgb.astype(np.int32).sum(axis=2).astype(np.float32)
    out = str(tmp_path / "out.flt")
    d = tmp_path / "pngs"
    d.mkdir()
    assert ske.run_cli(["--kernel", kernel, "--input", png, "--output", out, "--png-dir", str(d)]) == 0
    got = np.fromfile(out, np.float32).reshape(n, n)
    ref, _ = oracle.execute(oracle_id, grey, nthreads=1)
    if oracle_id == oracle.IDENTITYIMG:
        assert np.array_equal(got, ref)
    else:
        assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    for name in ("salida1.png", "salida2.png"):
        assert os.path.getsize(d / name) > 0
    # salida1 is the grey level / 3 (src/kernelRADN.h:385)
    s1 = np.asarray(Image.open(d / "salida1.png"))
    assert np.array_equal(s1, (grey / 3).astype(np.uint8))
