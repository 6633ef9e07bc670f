"""Standard parallel-beam Radon projector and filtered back-projection
(ske_fbp_* of include/skewengine.h) on the GPU."""
import ctypes

import numpy as np

from ._lib import check, lib

IMAGE, SINO, FILTERED, RECON = 0, 1, 2, 3


class FBP:
    """Projector / FBP of n x n images, n_angles x n_det sinograms.

    Buffers live on the device; upload/download move batches of images or
    sinograms (float32, C order)."""

    def __init__(self, n=512, n_angles=180, n_det=725, max_batch=1, angles=None, device_id=0):
        self.n, self.n_angles, self.n_det, self.max_batch = n, n_angles, n_det, max_batch
        self._h = ctypes.c_void_p()
        ang = None
        if angles is not None:
            angles = np.ascontiguousarray(angles, np.float32)
            ang = angles.ctypes.data
        check(lib().ske_fbp_create(n, n_angles, n_det, ang, max_batch, device_id, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            check(lib().ske_fbp_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, which, batch):
        if which in (IMAGE, RECON):
            return (batch, self.n, self.n)
        return (batch, self.n_angles, self.n_det)

    def upload(self, which, arr):
        arr = np.ascontiguousarray(arr, np.float32)
        batch = arr.shape[0] if arr.ndim == 3 else 1
        assert arr.size == int(np.prod(self._shape(which, batch)))
        check(lib().ske_fbp_upload(self._h, which, batch, arr.ctypes.data))
        return batch

    def download(self, which, batch=1):
        out = np.empty(self._shape(which, batch), np.float32)
        check(lib().ske_fbp_download(self._h, which, batch, out.ctypes.data))
        return out

    def project(self, batch=1):
        check(lib().ske_fbp_project(self._h, batch))

    def filter(self, batch=1):
        check(lib().ske_fbp_filter(self._h, batch))

    def backproject(self, batch=1):
        check(lib().ske_fbp_backproject(self._h, batch))

    def reconstruct(self, batch=1):
        check(lib().ske_fbp_reconstruct(self._h, batch))

    def set_taps(self, taps=None):
        if taps is None:
            check(lib().ske_fbp_set_taps(self._h, None))
        else:
            taps = np.ascontiguousarray(taps, np.float32)
            assert taps.size == 2 * self.n_det - 1
            check(lib().ske_fbp_set_taps(self._h, taps.ctypes.data))

    def taps(self):
        out = np.empty(2 * self.n_det - 1, np.float32)
        check(lib().ske_fbp_get_taps(self._h, out.ctypes.data))
        return out

    def synchronize(self):
        check(lib().ske_fbp_synchronize(self._h))

    def set_profile(self, on=True):
        check(lib().ske_fbp_set_option(self._h, 3, 1 if on else 0))

    def profile(self, name):
        ms = ctypes.c_double(0)
        cnt = ctypes.c_int(0)
        check(lib().ske_fbp_profile_get(self._h, name.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    def profile_reset(self):
        check(lib().ske_fbp_profile_reset(self._h))

    # conveniences -------------------------------------------------------------
    def radon(self, images):
        images = np.asarray(images, np.float32)
        single = images.ndim == 2
        b = self.upload(IMAGE, images[None] if single else images)
        self.project(b)
        out = self.download(SINO, b)
        return out[0] if single else out

    def fbp(self, sinograms):
        sinograms = np.asarray(sinograms, np.float32)
        single = sinograms.ndim == 2
        b = self.upload(SINO, sinograms[None] if single else sinograms)
        self.reconstruct(b)
        out = self.download(RECON, b)
        return out[0] if single else out
