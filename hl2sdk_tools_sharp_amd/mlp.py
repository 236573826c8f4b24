"""The learned warm start for the bounce: host side of libvradhip's network.

A per-patch MLP (25 -> 128 -> 128 -> 128 -> 3, GELU) predicts each patch's
converged bounced light from its position, normal, direct light and Fourier
features of its position, as max(gain * direct + scale * y, 0): the network
output y corrects a start proportional to the patch's own direct light (the
gain is the bounced-to-direct ratio of the training maps, the scale their
mean bounced light).  Training, inference, the features and the data
set all run on the GPU (BF16 MFMA GEMMs, FP32 accumulation, Adam); this class
only drives them.  The bounce then iterates from the prediction to the same
residual (``HipBounceStage.bounce(..., from_prediction=True)``), so the network
moves the starting point, never the answer.  The network has no counterpart
in the reference.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _lib
from .tracer import HipRayTracer

N_PARAMS = _lib.VH_MLP_PARAMS
N_INPUTS = _lib.VH_MLP_INPUTS
LAYERS = [(N_INPUTS, 128), (128, 128), (128, 128), (128, 3)]
FOURIER_SEED = 7
FOURIER_LENGTH = 1024.0
LIGHT_FLOOR = 1e-7  # the direct light enters the network as log(max(x, 0) + floor)


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def fourier_matrix(seed: int = FOURIER_SEED) -> np.ndarray:
    """The 8 x 3 frequency matrix of the position features (standard normal)."""
    return np.random.default_rng(seed).standard_normal((8, 3)).astype(np.float32)


def init_params(seed: int = 0, out_bias=None, out_scale: float = 0.1) -> np.ndarray:
    """He-normal hidden weights, a scaled-down output layer, zero biases
    (``out_bias``, 3 values, seeds the output bias); packed per layer as
    W [in][out] then b."""
    rng = np.random.default_rng(seed)
    parts = []
    for li, (i, o) in enumerate(LAYERS):
        W = rng.standard_normal((i, o)) * np.sqrt(2.0 / i)
        b = np.zeros(o)
        if li == len(LAYERS) - 1:
            W *= out_scale
            if out_bias is not None:
                b[:] = out_bias
        parts += [W.reshape(-1), b]
    flat = np.concatenate(parts).astype(np.float32)
    assert flat.shape == (N_PARAMS,)
    return flat


class WarmStartMLP:
    """The network on one tracer context (the context of the bounce stage it
    warm-starts)."""

    def __init__(self, tracer: HipRayTracer, params: Optional[np.ndarray] = None, seed: int = 0,
                 gain: float = 1.0, scale: float = 1.0, length: float = FOURIER_LENGTH,
                 floor: float = LIGHT_FLOOR):
        self.tracer = tracer
        self.opts = _lib.VhMlpOpts(float(length), float(floor), float(gain), float(scale))
        self.freq = fourier_matrix()
        p = init_params(seed) if params is None else np.ascontiguousarray(params, np.float32)
        if p.shape != (N_PARAMS,):
            raise ValueError(f"expected {N_PARAMS} parameters, got {p.shape}")
        self._call(_lib.lib().vh_mlp_init, _p(p), _p(self.freq), ctypes.byref(self.opts))

    def _call(self, fn, *args):
        HipRayTracer._raise(fn(self.tracer.handle, *args))

    def params(self) -> np.ndarray:
        out = np.zeros(N_PARAMS, np.float32)
        self._call(_lib.lib().vh_mlp_params, _p(out))
        return out

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_INPUTS)
        y = np.zeros((x.shape[0], 3), np.float32)
        self._call(_lib.lib().vh_mlp_forward, _p(x), x.shape[0], _p(y))
        return y

    def grad(self, x: np.ndarray, t: np.ndarray):
        """(loss, gradient) of the mean squared error over the rows."""
        x = np.ascontiguousarray(x, np.float32).reshape(-1, N_INPUTS)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, 3)
        g = np.zeros(N_PARAMS, np.float32)
        loss = ctypes.c_float()
        self._call(_lib.lib().vh_mlp_grad, _p(x), _p(t), x.shape[0], _p(g), ctypes.byref(loss))
        return loss.value, g

    def adam(self, grad: np.ndarray, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999,
             eps: float = 1e-8) -> None:
        g = np.ascontiguousarray(grad, np.float32)
        self._call(_lib.lib().vh_mlp_adam, _p(g), lr, beta1, beta2, eps)

    def features(self, emit0=None, npatch: Optional[int] = None) -> np.ndarray:
        """The inputs of the loaded patches (npatch x 25)."""
        e = None if emit0 is None else np.ascontiguousarray(emit0, np.float32).reshape(-1, 3)
        n = e.shape[0] if e is not None else int(npatch)
        x = np.zeros((n, N_INPUTS), np.float32)
        self._call(_lib.lib().vh_mlp_features, _p(e), _p(x))
        return x

    def dataset_clear(self) -> None:
        self._call(_lib.lib().vh_mlp_dataset, _lib.VH_MLP_DATA_CLEAR, None)

    def dataset_add(self) -> int:
        """Adds the last bounce's leaf patches (direct light in, bounced light
        out); returns the data set's size."""
        n = ctypes.c_int64()
        self._call(_lib.lib().vh_mlp_dataset, _lib.VH_MLP_DATA_ADD, ctypes.byref(n))
        return n.value

    def train(self, epochs: int = 10, batch: int = 4096, lr: float = 1e-3, seed: int = 0,
              use_graph: bool = True, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8,
              fused: bool = True) -> np.ndarray:
        """Epochs over the data set; returns each epoch's mean batch loss.
        ``fused`` trains with the two-launch step (rows, then weights and
        Adam), which gives the per-layer kernels' parameters bit for bit."""
        o = _lib.VhMlpTrainOpts(int(epochs), int(batch), int(seed), 1 if use_graph else 0, float(lr),
                                float(beta1), float(beta2), float(eps), 1 if fused else 0)
        losses = np.zeros(epochs, np.float32)
        self._call(_lib.lib().vh_mlp_train, ctypes.byref(o), _p(losses))
        return losses

    def predict(self, emit0=None, npatch: Optional[int] = None, read: bool = True):
        """The predicted bounced light of the loaded patches, kept on the
        device for the next from-prediction bounce; (npatch, 3) when read."""
        e = None if emit0 is None else np.ascontiguousarray(emit0, np.float32).reshape(-1, 3)
        out = None
        if read:
            n = e.shape[0] if e is not None else int(npatch)
            out = np.zeros((n, 3), np.float32)
        self._call(_lib.lib().vh_mlp_predict, _p(e), _p(out))
        return out


class Prior:
    """Accumulates, over training maps, what the network's prior needs: the
    bounced-to-direct light ratio (gain) and the mean bounced light (scale),
    over leaf patches that are not sky."""

    def __init__(self):
        self.direct = 0.0
        self.bounced = 0.0
        self.rows = 0

    def add(self, g, direct, bounced) -> None:
        m = g.leaf & ~g.sky.astype(bool)
        self.direct += float(np.asarray(direct, np.float64)[m].sum())
        self.bounced += float(np.asarray(bounced, np.float64)[m].sum())
        self.rows += int(m.sum()) * 3

    @property
    def gain(self) -> float:
        return self.bounced / self.direct if self.direct > 0 else 0.0

    @property
    def scale(self) -> float:
        return self.bounced / self.rows if self.rows and self.bounced > 0 else 1.0


def bounce_map(tracer: HipRayTracer, g, tol: float = 1e-6, max_iter: int = 300, weights: str = "f32"):
    """One seeded map on the tracer's context: its scene, patches, transfers,
    point-light direct light and a bounce to ``tol``.  Returns (stage, direct
    light, result); the bounce's state stays on the device for
    ``WarmStartMLP.dataset_add``."""
    from . import rooms
    from .bounce import HipBounceStage
    from .tracer import unpack_bits

    tracer.load_scene(g.casters)
    st = HipBounceStage(tracer, g)
    st.load_patches()
    st.build_transfers()
    bits = tracer.trace_point_lights(rooms.sample_points(g), g.lights)
    e0 = rooms.direct_light(g, unpack_bits(bits, g.npatch * len(g.lights)))
    r = st.bounce(e0, tol=tol, max_iter=max_iter, weights=weights)
    return st, e0, r
