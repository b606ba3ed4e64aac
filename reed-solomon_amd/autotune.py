"""Autotuner and learned cost model for route / strip selection.

Candidates for one apply shape (field, k inputs, m outputs, shard length L)
are a device route -- "table" (the matrix route with runtime coefficients:
the table kernel below 128 KiB strips, the bit-sliced planes kernel above),
"bitsliced" (the FFT route: the generated bit-sliced LCH program of a GF(2^8)
shape that has one, else the staged LCH FFT program on the LDS interpreter --
GF(2^8) on byte planes, src/fft.rs:393-592, or the tower) or "mfma"
(bit-sliced int8 MFMA, GF(2^8), k <= 32) -- times a
column strip size (one launch per strip, 0 = whole shard), the GPU analogue of
the reference's hand-tuned STRIP_BYTES (src/reedsolomon16.rs:796-804) and of
its FFT-vs-matrix routing rule (src/fft16.rs:85-106, src/reedsolomon.rs:204-234).

* ``exhaustive(shape)`` measures every candidate with ``trs_tune_apply``
  (kernel time on synthetic device data, HIP events).  Applies larger than
  ``cap_bytes`` are timed on a prefix window of the columns and scaled by
  L / window: kernel time is linear in the column count at those sizes.
* ``CostModel`` wraps the library's GPU MLP (12 features -> 64 -> 64 -> 1,
  log time), trained on the exhaustive timings.
* ``Tuner.choose(shape)`` ranks all candidates with the model and measures
  only the top ``top_n``; the exhaustive optimum judges its regret.
* ``heuristic(shape)`` is the method without the model: the reference's
  routing rules, whole shard -- GF(2^8): the generated program when the shape
  has one, the staged program for the shapes src/reedsolomon.rs:204-234
  takes; GF(2^16): the FFT when 9 * multiplies < 2 k m (src/fft16.rs:105).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import lib

ROUTES = ("table", "bitsliced", "mfma")
STRIPS = (0, 8 << 10, 16 << 10, 32 << 10, 64 << 10, 256 << 10, 1 << 20, 4 << 20)
GENERATED = {(7, 13), (10, 10), (14, 14), (16, 16), (18, 6), (32, 32), (10, 4)}
FEATURES = 12


@dataclass(frozen=True)
class Shape:
    field: int  # 8 or 16
    k: int
    m: int
    length: int  # bytes per shard


Candidate = Tuple[str, int]  # (route, strip bytes)


def sample_shapes(n: int, seed: int) -> List[Shape]:
    """The cost-model workload: field in {8, 16}; k uniform in [2, 256]
    (GF(2^8), k+m <= 256) or [2, 1024] (GF(2^16)); m uniform in [1, min(k, 256)];
    shard length log-uniform in [1 KiB, 4 MiB] (even for GF(2^16))."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        field = 8 if rng.random() < 0.5 else 16
        if field == 8:
            k = int(rng.integers(2, 257))
            m = int(rng.integers(1, min(k, 256) + 1))
            if k + m > 256:
                continue
        else:
            k = int(rng.integers(2, 1025))
            m = int(rng.integers(1, min(k, 256) + 1))
        length = int(round(math.exp(rng.uniform(math.log(1024), math.log(4 << 20)))))
        if field == 16:
            length -= length & 1
        out.append(Shape(field, k, m, max(2, length)))
    return out


def _tower_fft(s: Shape) -> bool:
    return 2 <= s.k <= 1024 and s.k + s.m <= 4096


def _byte_staged(s: Shape) -> bool:
    return s.field == 8 and (s.k, s.m) not in GENERATED


def candidates(s: Shape) -> List[Candidate]:
    routes = ["table"]
    if (s.field == 8 and (s.k, s.m) in GENERATED) or (s.field == 16 and _tower_fft(s)) \
            or _byte_staged(s):
        routes.append("bitsliced")
    if s.field == 8 and s.k <= 32:
        routes.append("mfma")
    return [(r, st) for r in routes for st in STRIPS]


def features(s: Shape, c: Candidate) -> np.ndarray:
    route, strip = c
    eff = s.length if strip == 0 or strip >= s.length else strip
    nstrips = math.ceil(s.length / eff)
    return np.array([
        1.0 if s.field == 16 else 0.0,
        math.log2(s.k),
        math.log2(s.m),
        math.log2(s.length),
        math.log2(s.k * s.m),
        math.log2((s.k + s.m) * s.length),
        1.0 if route == "table" else 0.0,
        1.0 if route == "bitsliced" else 0.0,
        1.0 if route == "mfma" else 0.0,
        math.log2(eff),
        math.log2(nstrips),
        s.k / (s.k + s.m),
    ], np.float32)


def measure(s: Shape, c: Candidate, cap_bytes: int = 32 << 20, reps: int = 2) -> Optional[float]:
    """Milliseconds of one apply of candidate c on shape s, or None if the
    route has no kernel for the shape."""
    route, strip = c
    window = max(64, min(s.length, cap_bytes // (s.k + s.m)))
    if window < s.length:
        window -= window % 64  # a column window keeps the shard's 16-byte vector accesses
    window -= window % (2 if s.field == 16 else 1)
    window = max(2, window)
    ms = ctypes.c_double(0.0)
    rc = lib.trs_tune_apply(s.field, s.k, s.m, window, ROUTES.index(route), strip, reps,
                            ctypes.byref(ms))
    if rc == -4:
        return None
    if rc != 0:
        raise RuntimeError(f"trs_tune_apply failed ({rc}) for {s} {c}")
    return ms.value * (s.length / window)


def exhaustive(s: Shape, **kw) -> Dict[Candidate, float]:
    out = {}
    for c in candidates(s):
        t = measure(s, c, **kw)
        if t is not None:
            out[c] = t
    return out


def heuristic(s: Shape) -> Candidate:
    if s.field == 8 and (s.k, s.m) in GENERATED:
        return ("bitsliced", 0)
    if _byte_staged(s) and s.length >= 32:
        from . import fft8_program_info

        info = fft8_program_info(s.k, s.m)
        if info is not None and info["reference_route"]:
            return ("bitsliced", 0)
    if s.field == 16 and _tower_fft(s):
        from . import fft16_program_info

        info = fft16_program_info(s.k, s.m)
        if info is not None and info["profitable"]:
            return ("bitsliced", 0)
    return ("table", 0)


class CostModel:
    """GPU MLP over standardised features, predicting standardised log time."""

    def __init__(self, seed: int = 0):
        h = ctypes.c_void_p()
        rc = lib.trs_mlp_new(seed, ctypes.byref(h))
        if rc:
            raise RuntimeError(f"trs_mlp_new failed ({rc})")
        self._h = h
        self.mu = np.zeros(FEATURES, np.float32)
        self.sd = np.ones(FEATURES, np.float32)
        self.tmu, self.tsd = 0.0, 1.0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and lib is not None:
            try:
                lib.trs_mlp_free(h)
            except Exception:
                pass
            self._h = None

    @staticmethod
    def _fp(a: np.ndarray):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def params(self) -> np.ndarray:
        n = lib.trs_mlp_param_count()
        out = np.zeros(n, np.float32)
        assert lib.trs_mlp_get_params(self._h, self._fp(out), n) == 0
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert lib.trs_mlp_set_params(self._h, self._fp(p), p.size) == 0

    def raw_gradients(self, X: np.ndarray, T: np.ndarray):
        """Loss and gradient on already-standardised inputs (for checks)."""
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(lib.trs_mlp_param_count(), np.float32)
        loss = ctypes.c_float(0)
        rc = lib.trs_mlp_gradients(self._h, self._fp(X), self._fp(T), len(T), self._fp(g),
                                   ctypes.byref(loss))
        assert rc == 0, rc
        return loss.value, g

    def raw_predict(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        y = np.zeros(X.shape[0], np.float32)
        assert lib.trs_mlp_predict(self._h, self._fp(X), X.shape[0], self._fp(y)) == 0
        return y

    def raw_train(self, X: np.ndarray, T: np.ndarray, steps: int, lr: float) -> float:
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        loss = ctypes.c_float(0)
        rc = lib.trs_mlp_train(self._h, self._fp(X), self._fp(T), len(T), steps, lr,
                               ctypes.byref(loss))
        assert rc == 0, rc
        return loss.value

    def fit(self, X: np.ndarray, times_ms: np.ndarray, steps: int = 3000,
            lr: float = 3e-3) -> float:
        self.mu = X.mean(0).astype(np.float32)
        self.sd = (X.std(0) + 1e-6).astype(np.float32)
        t = np.log(times_ms)
        self.tmu, self.tsd = float(t.mean()), float(t.std() + 1e-9)
        return self.raw_train((X - self.mu) / self.sd, (t - self.tmu) / self.tsd, steps, lr)

    def predict_ms(self, X: np.ndarray) -> np.ndarray:
        y = self.raw_predict((X - self.mu) / self.sd)
        return np.exp(y * self.tsd + self.tmu)


class Tuner:
    """Cost-model-guided search: rank every candidate, measure the top few."""

    def __init__(self, model: CostModel, top_n: int = 3):
        self.model = model
        self.top_n = top_n

    def rank(self, s: Shape) -> List[Candidate]:
        cands = candidates(s)
        pred = self.model.predict_ms(np.stack([features(s, c) for c in cands]))
        return [cands[i] for i in np.argsort(pred)]

    def choose(self, s: Shape, measured: Optional[Dict[Candidate, float]] = None) -> Candidate:
        """Best of the top_n ranked candidates, by measurement (taken from
        ``measured`` when given, so evaluation does not re-time)."""
        best, best_t = None, math.inf
        tried = 0
        for c in self.rank(s):
            t = measured.get(c) if measured is not None else measure(s, c)
            if t is None:
                continue
            tried += 1
            if t < best_t:
                best, best_t = c, t
            if tried >= self.top_n:
                break
        return best


def regret(measured: Dict[Candidate, float], chosen: Candidate) -> float:
    """time(chosen) / time(best) - 1."""
    return measured[chosen] / min(measured.values()) - 1.0


def dataset(shapes: Sequence[Shape], timings: Sequence[Dict[Candidate, float]]):
    X, T = [], []
    for s, meas in zip(shapes, timings):
        for c, t in meas.items():
            X.append(features(s, c))
            T.append(t)
    return np.stack(X), np.array(T, np.float64)
