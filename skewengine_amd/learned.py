"""Learned sparse-view FBP (ske_learned_* of include/skewengine.h): a learnable
symmetric filter in place of Ram-Lak and a residual 4-layer CNN, trained on
the GPU by the library's own forward/backward kernels and Adam.

Typical use::

    with FBP(512, 60, 725, max_batch=16) as f, LearnedFBP(f) as m:
        m.train(steps=1250, batch=16)            # 20k phantom/sinogram pairs
        rec = m.reconstruct_sinograms(sinos)     # (B, 60, 725) -> (B, 512, 512)

Plain Ram-Lak FBP (FBP.fbp) is the fallback and the yardstick: `evaluate`
reports the PSNR of both on held-out phantoms.
"""
import ctypes
import os
import time

import numpy as np

from . import _lib
from ._lib import check, lib
from .fbp import IMAGE, RECON, SINO

TAPS = _lib.LEARNED_TAPS  # taps offset of the 16-channel layout (LearnedFBP.taps_offset in general)
R0, A1, A2, A3, DR0, FILTERED_GRAD = (_lib.LEARNED_R0, _lib.LEARNED_A1, _lib.LEARNED_A2, _lib.LEARNED_A3,
                                      _lib.LEARNED_DR0, _lib.LEARNED_FILTERED_GRAD)
WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")

# Training-data model of the sparse-view benchmark: 10 random ellipses,
# Poisson transmission noise at I0 = 1e5 counts with attenuation MU per unit
# line integral (image side = 2 units).
N_ELLIPSES = 10
I0 = 1e5
MU = 1.0
HELD_OUT_SEED = 1 << 40  # seeds at and above this are never used for training


def random_ellipses(seed, batch, n_ellipses=N_ELLIPSES):
    """(batch, n_ellipses, 6) ellipses (x0, y0, a, b, phi, rho) of the
    library's generator for `seed`."""
    out = np.empty((batch, n_ellipses, 6), np.float32)
    check(lib().ske_random_ellipses(seed, batch, n_ellipses, out.ctypes.data))
    return out


def psnr(rec, target):
    """Per-image 10 log10(range^2 / mse), range = max - min of the target."""
    rec = np.asarray(rec, np.float64)
    target = np.asarray(target, np.float64)
    if target.ndim == 2:
        rec, target = rec[None], target[None]
    rng = target.max(axis=(1, 2)) - target.min(axis=(1, 2))
    mse = ((rec - target) ** 2).mean(axis=(1, 2))
    return 10.0 * np.log10(rng * rng / mse)


class LearnedFBP:
    """Learned reconstruction attached to an FBP object (same geometry,
    stream and buffers: SINO in, RECON out, IMAGE the training target)."""

    def __init__(self, fbp, params=None, seed=1, channels=16):
        self.fbp = fbp
        self._h = ctypes.c_void_p()
        check(lib().ske_learned_create_channels(fbp._h, int(channels), ctypes.byref(self._h)))
        cnt = ctypes.c_int(0)
        check(lib().ske_learned_param_count(self._h, ctypes.byref(cnt)))
        self.param_count = cnt.value
        c, t = ctypes.c_int(0), ctypes.c_int(0)
        check(lib().ske_learned_layout(self._h, ctypes.byref(c), ctypes.byref(t)))
        self.channels, self.taps_offset = c.value, t.value
        if params is not None:
            self.set_params(params)
        elif seed != 1:
            self.init(seed)

    def close(self):
        if self._h:
            check(lib().ske_learned_destroy(self._h))
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

    # parameters ---------------------------------------------------------------
    def init(self, seed=1):
        check(lib().ske_learned_init(self._h, seed))

    def set_params(self, params):
        params = np.ascontiguousarray(params, np.float32)
        if params.size != self.param_count:
            raise ValueError("expected %d parameters, got %d" % (self.param_count, params.size))
        check(lib().ske_learned_set_params(self._h, params.ctypes.data))

    def params(self):
        out = np.empty(self.param_count, np.float32)
        check(lib().ske_learned_get_params(self._h, out.ctypes.data))
        return out

    def save(self, path):
        np.save(path, self.params())

    def load(self, path):
        self.set_params(np.load(path, allow_pickle=False))

    # compute ------------------------------------------------------------------
    def reconstruct(self, batch=1):
        check(lib().ske_learned_reconstruct(self._h, batch))

    def gradient(self, batch=1):
        loss = ctypes.c_double(0)
        grad = np.empty(self.param_count, np.float32)
        check(lib().ske_learned_gradient(self._h, batch, ctypes.byref(loss), grad.ctypes.data))
        return loss.value, grad

    def train_step(self, batch, lr, lr_taps, want_loss=True):
        loss = ctypes.c_double(0)
        check(lib().ske_learned_train_step(self._h, batch, lr, lr_taps, ctypes.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def download(self, which, batch=1):
        n = self.fbp.n
        if which in (A1, A2, A3):
            out = np.empty((batch, n, n, self.channels), np.uint16)
        elif which == FILTERED_GRAD:
            out = np.empty((batch, self.fbp.n_angles, self.fbp.n_det), np.float32)
        else:
            out = np.empty((batch, n, n), np.float32)
        check(lib().ske_learned_download(self._h, which, batch, out.ctypes.data))
        return out

    def render(self, ellipses):
        ellipses = np.ascontiguousarray(ellipses, np.float32)
        batch, nell = ellipses.shape[0], ellipses.shape[1]
        check(lib().ske_learned_render(self._h, batch, nell, ellipses.ctypes.data))
        return batch

    def make_batch(self, batch, seed, i0=I0, mu=MU, n_ellipses=N_ELLIPSES):
        check(lib().ske_learned_make_batch(self._h, batch, n_ellipses, seed, i0, mu))

    def reconstruct_sinograms(self, sinograms):
        sinograms = np.asarray(sinograms, np.float32)
        single = sinograms.ndim == 2
        b = self.fbp.upload(SINO, sinograms[None] if single else sinograms)
        self.reconstruct(b)
        out = self.fbp.download(RECON, b)
        return out[0] if single else out

    # training / evaluation -----------------------------------------------------
    def train(self, steps, batch, lr=1e-3, lr_taps=1e-3, seed0=0, i0=I0, mu=MU, log_every=0, decay_from=0.5):
        """Adam on fresh phantom/sinogram pairs each step (seeds seed0 + step);
        the learning rates fall linearly to 10% over the last (1 - decay_from)
        of the steps.  Returns the loss history (every step)."""
        hist = []
        t0 = time.perf_counter()
        for s in range(steps):
            frac = s / max(1, steps)
            scale = 1.0 if frac < decay_from else 1.0 - 0.9 * (frac - decay_from) / (1.0 - decay_from)
            self.make_batch(batch, seed0 + s, i0, mu)
            hist.append(self.train_step(batch, lr * scale, lr_taps * scale))
            if log_every and (s + 1) % log_every == 0:
                print("step %d/%d loss %.4e (%.1f s)" % (s + 1, steps, np.mean(hist[-log_every:]),
                                                         time.perf_counter() - t0), flush=True)
        return np.array(hist)

    def evaluate(self, count, batch, seed0=HELD_OUT_SEED, i0=I0, mu=MU):
        """PSNR of plain Ram-Lak FBP and of the learned reconstruction on
        `count` held-out phantoms: (psnr_fbp, psnr_learned) arrays."""
        p_fbp, p_l = [], []
        done = 0
        while done < count:
            b = min(batch, count - done)
            self.make_batch(b, seed0 + done, i0, mu)
            target = self.fbp.download(IMAGE, b)
            self.fbp.reconstruct(b)
            p_fbp.append(psnr(self.fbp.download(RECON, b), target))
            self.reconstruct(b)
            p_l.append(psnr(self.fbp.download(RECON, b), target))
            done += b
        return np.concatenate(p_fbp), np.concatenate(p_l)


def default_weights(n=512, n_angles=60, n_det=725, channels=16):
    """Path of the committed trained parameters for this geometry and
    channel count, or None."""
    suffix = "" if channels == 16 else "_c%d" % channels
    p = os.path.join(WEIGHTS_DIR, "learned_fbp_n%d_a%d_d%d%s.npy" % (n, n_angles, n_det, suffix))
    return p if os.path.exists(p) else None


def main(argv=None):
    """Train the learned sparse-view FBP on the GPU and save its parameters:

        python -m skewengine_amd.learned --steps 1250 --batch 16 --out weights.npy
    """
    import argparse
    from .fbp import FBP
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--angles", type=int, default=60)
    ap.add_argument("--det", type=int, default=725)
    ap.add_argument("--channels", type=int, default=16, choices=(16, 32))
    ap.add_argument("--steps", type=int, default=1250)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-taps", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--init", default=None, help="start from these parameters (.npy)")
    ap.add_argument("--eval", type=int, default=32, help="held-out phantoms for the PSNR report")
    ap.add_argument("--log-every", type=int, default=50)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    with FBP(a.n, a.angles, a.det, max_batch=a.batch) as f, LearnedFBP(f, seed=a.seed, channels=a.channels) as m:
        if a.init:
            m.load(a.init)
        t0 = time.perf_counter()
        hist = m.train(a.steps, a.batch, a.lr, a.lr_taps, log_every=a.log_every)
        dt = time.perf_counter() - t0
        print("trained %d steps x %d pairs in %.1f s (%.2f ms/step), final loss %.4e"
              % (a.steps, a.batch, dt, 1e3 * dt / max(1, a.steps), np.mean(hist[-20:])), flush=True)
        if a.out:
            m.save(a.out)
        if a.eval:
            p_fbp, p_l = m.evaluate(a.eval, a.batch)
            print("held-out PSNR over %d phantoms: plain FBP %.2f dB, learned %.2f dB, gain %.2f dB (min %.2f)"
                  % (a.eval, p_fbp.mean(), p_l.mean(), (p_l - p_fbp).mean(), (p_l - p_fbp).min()), flush=True)


if __name__ == "__main__":
    main()
