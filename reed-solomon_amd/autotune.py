"""Autotuner and learned cost model for kernel selection.

Candidates for one encode-shaped apply (field, k inputs, m outputs, shard
length L) are the library's kernel variants (include/trs_gpu.h): the table
tile, the bit-sliced planes tile with 4 or 8 rows per workgroup, int8 MFMA
(GF(2^8), k <= 32), and the shape's FFT program -- the generated bit-sliced
program of a GF(2^8) shape that has one, else the staged LCH program on the
LDS interpreter (GF(2^8) byte planes, src/fft.rs:393-592, or the tower,
src/fft16.rs:357-376) sized for 16 / 32 / 64 KiB of LDS per workgroup.
Launch parameters that come out the same are one candidate.  These replace
the reference's hand-tuned constants: MULT_MARGIN (src/reedsolomon.rs:226-232),
STRIP_BYTES (src/reedsolomon16.rs:796-804) and its FFT-vs-matrix rule
(src/fft16.rs:85-106).  The library computes the candidate set, the model's
features and its own rules, so training here and routing in the library see
the same numbers.

* ``exhaustive(shape)`` times every candidate with ``trs_tune_apply`` (HIP
  events, synthetic device data).  Applies larger than ``cap_bytes`` are
  timed on a prefix window of the columns and scaled by L / window: kernel
  time is linear in the column count at those sizes.
* ``CostModel`` wraps the library's GPU MLP (12 features -> 64 -> 64 -> 1,
  log time), trained on the exhaustive timings; ``install()`` hands it to the
  library's router, which then ranks the candidates of every automatic call
  it has not seen, times the top few on that call and caches the winner.
* ``Tuner.choose(shape)`` ranks the candidates with the model and measures
  only the top ``top_n``; the exhaustive optimum judges its regret.
* ``rule(shape)`` is the codec's own GPU-calibrated rule (no model);
  ``heuristic(shape)`` the reference's CPU routing -- GF(2^8): the generated
  program when the shape has one, the staged program for the shapes
  src/reedsolomon.rs:204-234 takes; GF(2^16): the FFT when
  9 * multiplies < 2 k m (src/fft16.rs:105) -- with the codec's matrix kernel
  otherwise.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np

from ._lib import lib

VARIANTS = ("table", "planes4", "planes8", "mfma", "program16k", "program32k", "program64k")
PROGRAM_DEFAULT = {8: 5, 16: 6}  # the FFT program at the codec's default LDS target
GENERATED = {(7, 13), (10, 10), (14, 14), (16, 16), (18, 6), (32, 32), (10, 4)}
FEATURES = 12


@dataclass(frozen=True)
class Shape:
    field: int  # 8 or 16
    k: int
    m: int
    length: int  # bytes per shard


Candidate = int  # variant id, an index into VARIANTS


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


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


def candidates(s: Shape) -> List[Candidate]:
    mask = ctypes.c_uint32(0)
    _check(lib.trs_tune_feasible(s.field, s.k, s.m, s.length, ctypes.byref(mask)),
           "trs_tune_feasible")
    return [v for v in range(len(VARIANTS)) if (mask.value >> v) & 1]


def features(s: Shape, c: Candidate) -> np.ndarray:
    out = np.zeros(FEATURES, np.float32)
    _check(lib.trs_tune_features(s.field, s.k, s.m, s.length, c,
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), FEATURES),
           "trs_tune_features")
    return out


def measure(s: Shape, c: Candidate, cap_bytes: int = 64 << 20, min_window: int = 1 << 20,
            reps: int = 2, limit_ms: float = 0.0) -> Optional[float]:
    """Milliseconds of one apply of candidate c on shape s, or None if the
    variant has no kernel for the shape.  A long apply is timed on a prefix
    window of its columns -- at least min_window bytes per shard, so that every
    kernel still gets a full-size grid, else cap_bytes over all shards -- and
    scaled by L / window.  limit_ms: report the warm-up alone when it is
    already slower."""
    window = min(s.length, max(min_window, cap_bytes // (s.k + s.m)))
    if window < s.length:
        window -= window % 64  # a column window keeps the shard's 16-byte vector accesses
    window -= window % (2 if s.field == 16 else 1)
    window = max(2, window)
    ms = ctypes.c_double(0.0)
    rc = lib.trs_tune_apply(s.field, s.k, s.m, window, c, reps, limit_ms * window / s.length,
                            ctypes.byref(ms))
    if rc == -4:
        return None
    _check(rc, f"trs_tune_apply {s} {VARIANTS[c]}")
    return ms.value * (s.length / window)


def exhaustive(s: Shape, **kw) -> Dict[Candidate, float]:
    """Every candidate's time; the FFT programs (usually fastest) first, and a
    candidate whose warm-up is already 10x the best so far keeps that time."""
    out: Dict[Candidate, float] = {}
    for c in sorted(candidates(s), key=lambda v: -v):
        best = min(out.values()) if out else 0.0
        t = measure(s, c, limit_ms=10.0 * best, **kw)
        if t is not None:
            out[c] = t
    return out


def rule(s: Shape) -> Candidate:
    """The codec's own GPU-calibrated choice for an automatic encode."""
    v = ctypes.c_int(-1)
    _check(lib.trs_tune_rule(s.field, s.k, s.m, s.length, ctypes.byref(v)), "trs_tune_rule")
    return v.value


def heuristic(s: Shape) -> Candidate:
    """The reference's CPU routing, with the codec's matrix kernel for its
    matrix route."""
    fft = False
    if s.field == 8 and (s.k, s.m) in GENERATED:
        fft = True
    elif s.field == 8 and s.length >= 32:
        from . import fft8_program_info

        info = fft8_program_info(s.k, s.m)
        fft = info is not None and bool(info["reference_route"])
    elif s.field == 16 and 2 <= s.k <= 1024 and s.k + s.m <= 4096:
        from . import fft16_program_info

        info = fft16_program_info(s.k, s.m)
        fft = info is not None and bool(info["profitable"])
    cands = candidates(s)
    if fft and PROGRAM_DEFAULT[s.field] in cands:
        return PROGRAM_DEFAULT[s.field]
    # the codec's matrix kernel by length (its rule without the FFT program)
    if s.field == 16:
        return 0
    if 3 in cands and s.length < (512 << 10):
        return 3
    if s.length < (128 << 10):
        return 0
    wg8 = math.ceil(s.length / 8192) * math.ceil(s.m / 8)
    return 1 if s.m <= 4 or wg8 <= 640 else 2


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

    def install(self, top_n: int = 3) -> None:
        """Hands the trained model to the library's router (trs_tuner_install)."""
        p = self.params()
        norm = np.ascontiguousarray(np.concatenate([self.mu, self.sd]), np.float32)
        _check(lib.trs_tuner_install(self._fp(p), p.size, self._fp(norm), norm.size, top_n),
               "trs_tuner_install")


def uninstall() -> None:
    """Back to the codec's rules (trs_tuner_remove)."""
    _check(lib.trs_tuner_remove(), "trs_tuner_remove")


def routed_choice(field: int, op: int, k: int, nout: int, length: int) -> int:
    """The router's cached variant for a key (op 0 encode, 1 reconstruct), -1 if none."""
    v = ctypes.c_int(-1)
    _check(lib.trs_tuner_lookup(field, op, k, nout, length, ctypes.byref(v)), "trs_tuner_lookup")
    return v.value


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
