"""ske_execute_noalloc (no device allocation: every buffer beyond the
caller's input and output lives in the library's static scratch array)
against ske_execute, which allocates an engine per call.  Bit-exact for
every operator: both run the same kernels over the same sector order; the
no-allocation path only splits the sectors into other chunks.  Two calls in
a row on the same array, the second on a smaller grid, must both match."""
import numpy as np
import pytest

import skewengine_amd as ske
from skewengine_amd import _lib, synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _noalloc(kernel, grid, s0=0, s1=180, local=0, povh=0.15):
    n = grid.shape[0]
    d_in = torch.from_numpy(np.ascontiguousarray(grid, np.float32)).cuda()
    d_out = torch.full((n, n), 7.0, dtype=torch.float32, device="cuda")  # stale values must be overwritten
    torch.cuda.synchronize()
    rc = _lib.lib().ske_execute_noalloc(kernel, n, d_in.data_ptr(), d_out.data_ptr(), povh, s0, s1, local, None)
    _lib.check(rc)
    return d_out.cpu().numpy()


def _engine(kernel, grid, s0=0, s1=180, local=0, povh=0.15):
    with ske.SkewEngine(grid) as e:
        if local:
            e.set_option(_lib.OPT_RADON_LOCAL, 1)
        return e.run(kernel, s0, s1, povh)


def test_viewshed_largest_then_smaller_grid():
    big = synth.dem(1024, 0.8, 31, 1500.0)
    small = synth.dem(200, 0.85, 32, 1500.0)
    print("noalloc: viewshed 1024^2 then 200^2", flush=True)
    a = _noalloc(ske.KERNEL_SDEM, big)
    b = _noalloc(ske.KERNEL_SDEM, small)
    assert np.array_equal(a, ske.execute(ske.KERNEL_SDEM, big))
    assert np.array_equal(b, ske.execute(ske.KERNEL_SDEM, small))


@pytest.mark.parametrize("kernel,n,local", [
    (ske.KERNEL_CEPSTRUM, 300, 0), (ske.KERNEL_RADON, 256, 0), (ske.KERNEL_RADON, 96, 1),
    (ske.KERNEL_IDENTITYDEM, 97, 0), (ske.KERNEL_IDENTITYIMG, 1024, 0)])
def test_operators_two_calls_match_engine(kernel, n, local):
    grid = synth.blurred_texture(n, 0.6, seed=n)[0] if kernel != ske.KERNEL_IDENTITYDEM else synth.dem(n, 0.8, n, 1500.0)
    smaller = grid[: n // 2, : n // 2].copy()
    print("noalloc: operator %d n=%d local=%d" % (kernel, n, local), flush=True)
    first = _noalloc(kernel, grid, local=local)
    second = _noalloc(kernel, smaller, 10, 170, local=local)
    assert np.array_equal(first, _engine(kernel, grid, local=local))
    assert np.array_equal(second, _engine(kernel, smaller, 10, 170, local=local))


def test_rejects_sizes_beyond_the_array():
    L = _lib.lib()
    d = torch.zeros(1025 * 1025, dtype=torch.float32, device="cuda")
    assert L.ske_execute_noalloc(0, 1025, d.data_ptr(), d.data_ptr(), 0.15, 0, 180, 0, None) == _lib.SKE_OK - 1
    assert b"maximum" in L.ske_last_error()
    assert L.ske_execute_noalloc(0, 1, d.data_ptr(), d.data_ptr(), 0.15, 0, 180, 0, None) == _lib.SKE_OK - 1
    assert L.ske_execute_noalloc(0, 64, None, d.data_ptr(), 0.15, 0, 180, 0, None) == _lib.SKE_OK - 1
    assert L.ske_execute_noalloc(9, 64, d.data_ptr(), d.data_ptr(), 0.15, 0, 180, 0, None) == _lib.SKE_OK - 1
