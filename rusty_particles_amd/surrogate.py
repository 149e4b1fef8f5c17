"""JKR contact-radius surrogate: training and inference on the GPU.

The reference solves the JKR contact radius by Newton iteration for every
adhesive contact (src/physics.rs:132-219).  In scaled variables
alpha = a / s, d = delta / t (s = (R*^2 8 pi gamma / E*)^(1/3), t = s^2 / R*)
the equation is alpha^2 - sqrt(alpha) = d for every material and size, and
F / (pi gamma R*) = (32/3) alpha^3 - 8 alpha^(3/2).  A 4 -> 64 -> 64 -> 2 MLP
(BF16 MFMA, FP32 accumulate) maps (sqrt(d - d_min), ln R*, ln E*, ln gamma)
to (alpha, F / (pi gamma R*)); the engine polishes its alpha with FP64 Newton
steps and falls back to the exact solve outside the envelope or when the
residual check fails.  The envelope excludes 0 < d < 0.2, where the
reference's iteration starts left of the minimum of alpha^2 - sqrt(alpha)
and does not reach the larger root, and needs d_lo >= -0.40.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi, native

D_MIN = float(0.25 ** (4.0 / 3.0) - 0.25 ** (1.0 / 3.0))  # minimum of alpha^2 - sqrt(alpha): no root below
GAP_HI = 0.2       # 0 < d < GAP_HI: the exact solve
D_LO_MIN = -0.40   # lowest d served


def _lib():
    L = native.lib()
    if not getattr(L, "_surrogate_bound", False):
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        env = C.POINTER(abi.RpJkrEnvelope)
        L.rp_jkr_surrogate_train.argtypes = [env, C.c_uint64, C.c_size_t, C.c_int, C.c_int, C.c_float, fp, dp]
        L.rp_jkr_samples.argtypes = [env, C.c_uint64, C.c_size_t, fp, fp, dp]
        L.rp_jkr_envelope_scale.argtypes = [env, fp, fp, fp, fp]
        L.rp_mlp_infer.argtypes = [fp, fp, C.c_size_t, fp]
        L.rp_mlp_gradient.argtypes = [fp, fp, fp, C.c_size_t, fp, dp]
        L.rp_mfma_probe.argtypes = [fp, fp, fp]
        L.rp_surrogate_last_error.restype = C.c_char_p
        for n in ("rp_jkr_surrogate_train", "rp_jkr_samples", "rp_jkr_envelope_scale", "rp_mlp_infer",
                  "rp_mlp_gradient", "rp_mfma_probe"):
            getattr(L, n).restype = C.c_int32
        L._surrogate_bound = True
    return L


def _check(code):
    if code != abi.RP_OK:
        raise native.EngineError(code, _lib().rp_surrogate_last_error().decode())


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def envelope(d_lo, d_hi, rstar, estar, gamma) -> abi.RpJkrEnvelope:
    """Envelope from (lo, hi) pairs of R*, E*, gamma (a scalar means a fixed value)."""
    def pair(v):
        return (float(v), float(v)) if np.isscalar(v) else (float(v[0]), float(v[1]))
    e = abi.RpJkrEnvelope()
    e.d_lo, e.d_hi = float(d_lo), float(d_hi)
    e.rstar_lo, e.rstar_hi = pair(rstar)
    e.estar_lo, e.estar_hi = pair(estar)
    e.gamma_lo, e.gamma_hi = pair(gamma)
    return e


@dataclass
class JkrSurrogate:
    params: np.ndarray          # float32 [RP_MLP_PARAMS]
    env: abi.RpJkrEnvelope
    epoch_loss: np.ndarray | None = None

    def scale(self):
        in_lo, in_hi = np.zeros(4, np.float32), np.zeros(4, np.float32)
        out_lo, out_hi = np.zeros(2, np.float32), np.zeros(2, np.float32)
        _check(_lib().rp_jkr_envelope_scale(C.byref(self.env), _f(in_lo), _f(in_hi), _f(out_lo), _f(out_hi)))
        return in_lo, in_hi, out_lo, out_hi

    def infer_scaled(self, x_scaled: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x_scaled, dtype=np.float32).reshape(-1, 4)
        y = np.zeros((x.shape[0], 2), np.float32)
        _check(_lib().rp_mlp_infer(_f(np.ascontiguousarray(self.params, np.float32)), _f(x), x.shape[0], _f(y)))
        return y

    def predict(self, d, rstar, estar, gamma):
        """alpha and F / (pi gamma R*) for arrays of contacts."""
        in_lo, in_hi, out_lo, out_hi = self.scale()
        u = np.sqrt(np.asarray(d, float) - D_MIN)
        feats = np.stack([u, np.log(rstar) + 0 * u, np.log(estar) + 0 * u, np.log(gamma) + 0 * u], 1)
        x = 2.0 * (feats - in_lo) / (in_hi - in_lo) - 1.0
        y = self.infer_scaled(x).astype(np.float64)
        return out_lo + 0.5 * (y + 1.0) * (out_hi - out_lo)


def samples(env: abi.RpJkrEnvelope, n: int, seed: int = 1):
    """(x_scaled, t_scaled, raw) labelled on the GPU by the FP64 Newton solve."""
    x = np.zeros((n, 4), np.float32)
    t = np.zeros((n, 2), np.float32)
    raw = np.zeros((n, 8))
    _check(_lib().rp_jkr_samples(C.byref(env), seed, n, _f(x), _f(t), _d(raw)))
    return x, t, raw


def train(env: abi.RpJkrEnvelope, nsamples: int = 1_000_000, epochs: int = 40, batch: int = 4096, lr: float = 3e-3,
          seed: int = 7) -> JkrSurrogate:
    params = np.zeros(abi.RP_MLP_PARAMS, np.float32)
    losses = np.zeros(epochs)
    _check(_lib().rp_jkr_surrogate_train(C.byref(env), seed, nsamples, epochs, batch, lr, _f(params), _d(losses)))
    return JkrSurrogate(params, env, losses)


def gradient(params, x_scaled, t_scaled):
    x = np.ascontiguousarray(x_scaled, np.float32)
    t = np.ascontiguousarray(t_scaled, np.float32)
    g = np.zeros(abi.RP_MLP_PARAMS, np.float32)
    loss = C.c_double(0.0)
    _check(_lib().rp_mlp_gradient(_f(np.ascontiguousarray(params, np.float32)), _f(x), _f(t), x.shape[0], _f(g),
                                  C.byref(loss)))
    return loss.value, g


def mfma_probe(a, b):
    c = np.zeros((16, 16), np.float32)
    _check(_lib().rp_mfma_probe(_f(np.ascontiguousarray(a, np.float32)), _f(np.ascontiguousarray(b, np.float32)),
                                _f(c)))
    return c
