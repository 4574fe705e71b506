"""Learned elevation bracket: predictor objects, the native trainer and the
training-data pipeline (no counterpart in the reference).

The predictor is a 192 -> 256 -> 256 -> 24 GELU network.  Its input is the
per-cell descriptor computed on the device (24 azimuths x 8 log-spaced rings
of maximum terrain elevation angle, 0.25 .. 40 km, BF16); its output is the
horizon per azimuth in degrees.  ``search_mode=2`` starts each azimuth's
lattice search at the predicted elevation: the search climbs while the query
hits and descends while it misses, so where the hits along the azimuth are
monotone in elevation a wrong prediction costs queries, not accuracy.  Where
a ray can pass through a hole of the mesh they need not be, and there the
lattice search from any start can differ from the reference walk; with the
default ``certify=1`` the library runs the reference walk itself for every
cell with a mesh-boundary point within reach, so the learned search's
results equal the walk's everywhere (to the walk's FP64 rounding of its
carried elevation, ~1e-14 deg).

Training targets are horizons the method itself produces (the clip-aware
lattice search, ``search_mode=1``) on synthetic terrains from fresh seeds.
Forward, backward and Adam run on the device in hand-written BF16 MFMA
kernels (csrc/mlp.hip).
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib, synthetic
from ._lib import check
from .horizon_svf import ResidentDomain

N_IN, N_H1, N_H2, N_OUT = 192, 256, 256, 24
SHAPES = [("w1", (N_H1, N_IN)), ("b1", (N_H1,)), ("w2", (N_H2, N_H1)), ("b2", (N_H2,)),
          ("w3", (N_OUT, N_H2)), ("b3", (N_OUT,))]
WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
DEFAULT_WEIGHTS = os.path.join(WEIGHTS_DIR, "predictor_c2.npz")

_fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
_up = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))  # noqa: E731


def num_params() -> int:
    return int(_lib.lib().hz_predictor_num_params())


def unflatten(flat: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = flat[o:o + n].reshape(shp)
        o += n
    return out


def flatten(params: dict) -> np.ndarray:
    return np.concatenate([np.asarray(params[n], np.float32).ravel() for n, _ in SHAPES])


def save_params(path: str, flat: np.ndarray, **meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, **unflatten(np.asarray(flat, np.float32)),
             **{k: np.asarray(v) for k, v in meta.items()})


def load_params(path: str = DEFAULT_WEIGHTS) -> np.ndarray:
    z = np.load(path, allow_pickle=False)
    return flatten({n: z[n] for n, _ in SHAPES})


class Predictor:
    """Device copy of the network used by ``search_mode=2``."""

    def __init__(self, flat: np.ndarray):
        flat = np.ascontiguousarray(flat, np.float32)
        if flat.size != num_params():
            raise ValueError(f"expected {num_params()} parameters, got {flat.size}")
        h = ctypes.c_void_p()
        check(_lib.lib().hz_predictor_create(_fp(flat), ctypes.byref(h)))
        self._h = h
        self.params = flat

    @classmethod
    def from_file(cls, path: str = DEFAULT_WEIGHTS) -> "Predictor":
        return cls(load_params(path))

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().hz_predictor_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Trainer:
    """Native device trainer (Adam, MSE, cosine step decay)."""

    def __init__(self, seed: int = 0):
        h = ctypes.c_void_p()
        check(_lib.lib().hz_trainer_create(ctypes.c_ulonglong(seed), ctypes.byref(h)))
        self._h = h

    def set_data(self, X: np.ndarray, T: np.ndarray):
        X = np.ascontiguousarray(X, np.uint16)
        T = np.ascontiguousarray(T, np.float32)
        assert X.shape[1] == N_IN and T.shape == (X.shape[0], N_OUT)
        check(_lib.lib().hz_trainer_set_data(self._h, _up(X), _fp(T), X.shape[0]))

    def fit(self, epochs: int = 8, batch: int = 2048, lr: float = 2e-3) -> np.ndarray:
        loss = np.zeros(epochs, np.float32)
        check(_lib.lib().hz_trainer_fit(self._h, epochs, batch, ctypes.c_float(lr), _fp(loss)))
        return loss

    def gradients(self, X: np.ndarray, T: np.ndarray):
        X = np.ascontiguousarray(X, np.uint16)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(num_params(), np.float32)
        loss = ctypes.c_float()
        check(_lib.lib().hz_trainer_gradients(self._h, _up(X), _fp(T), X.shape[0], _fp(g),
                                              ctypes.byref(loss)))
        return g, float(loss.value)

    def predict(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.uint16)
        out = np.zeros((X.shape[0], N_OUT), np.float32)
        check(_lib.lib().hz_trainer_predict(self._h, _up(X), X.shape[0], _fp(out)))
        return out

    def get_weights(self) -> np.ndarray:
        w = np.zeros(num_params(), np.float32)
        check(_lib.lib().hz_trainer_get_weights(self._h, _fp(w)))
        return w

    def set_weights(self, flat: np.ndarray):
        flat = np.ascontiguousarray(flat, np.float32)
        check(_lib.lib().hz_trainer_set_weights(self._h, _fp(flat)))

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().hz_trainer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """float32 -> BF16 bit patterns (round to nearest even)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_to_f32(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, np.uint32) << 16).view(np.float32)


def domain_samples(dom, grid_type: int = 1, dist_search: float = 40000.0):
    """Descriptors (BF16 bits) and method-generated targets (clip-aware
    lattice search, deg) of every cell of a synthetic domain."""
    with ResidentDomain(*dom.args(), grid_type=grid_type, ray_org_elev=0.2) as rd:
        X = rd.descriptors()
        rd.run(24, dist_search, 1, 2, search_mode=1, certify=0)
        h, _ = rd.fetch()
    return X, np.ascontiguousarray(h.T, np.float32)


def train_predictor(seeds=(1001, 1002), config=synthetic.config2, epochs: int = 8,
                    batch: int = 2048, lr: float = 2e-3, max_cells: int = 1_000_000,
                    seed: int = 0, verbose: bool = True):
    """Train on fresh synthetic terrains; returns (flat params, epoch losses,
    number of training cells)."""
    Xs, Ts = [], []
    n = 0
    for s in seeds:
        X, T = domain_samples(config(seed=s))
        Xs.append(X)
        Ts.append(T)
        n += X.shape[0]
        if verbose:
            print(f"[train] seed {s}: {X.shape[0]} cells", flush=True)
        if n >= max_cells:
            break
    X = np.concatenate(Xs)[:max_cells]
    T = np.concatenate(Ts)[:max_cells]
    tr = Trainer(seed)
    tr.set_data(X, T)
    loss = tr.fit(epochs, batch, lr)
    if verbose:
        print(f"[train] {X.shape[0]} cells, loss per epoch {np.round(loss, 4).tolist()}",
              flush=True)
    w = tr.get_weights()
    tr.close()
    return w, loss, X.shape[0]


def main(argv=None):
    """Train the predictor on fresh C2-recipe terrains and write its weights:
    python -m horayzon_extpar_amd.learned --out <file.npz>"""
    import argparse
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--seeds", default="1001,1002")
    ap.add_argument("--cells", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=2e-3)
    a = ap.parse_args(argv)
    t0 = time.time()
    w, loss, n = train_predictor(seeds=tuple(int(x) for x in a.seeds.split(",")),
                                 epochs=a.epochs, batch=a.batch, lr=a.lr, max_cells=a.cells)
    save_params(a.out, w, loss=loss, cells=n, seeds=a.seeds, epochs=a.epochs)
    print(f"[train] wrote {a.out} ({n} cells, {time.time() - t0:.1f} s)", flush=True)


if __name__ == "__main__":
    main()
