"""Helpers shared by the GPU parity tests: device tensors from seeded bytes."""
import numpy as np

import oracle as O


def deterministic_data(length: int, seed: int) -> np.ndarray:
    """The reference tests' LCG byte stream (src/reedsolomon.rs:896-904)."""
    out = np.zeros(length, np.uint8)
    s = seed & 0xFFFFFFFF
    for i in range(length):
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        out[i] = (s >> 16) & 0xFF
    return out


def seeded(k: int, length: int, seed: int) -> np.ndarray:
    """k rows of the splitmix64 stream (fast; used for the larger cases)."""
    return O.fill(seed, k * length).reshape(k, length)


def to_dev(a: np.ndarray):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def rows_to_dev(rows: np.ndarray):
    return [to_dev(rows[i]) for i in range(rows.shape[0])]


def to_host(ts):
    return np.stack([t.cpu().numpy() for t in ts])
