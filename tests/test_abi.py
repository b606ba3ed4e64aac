"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
symbol include/trs_gpu.h declares, and validates arguments that need no device
(ReedSolomon::new errors, src/reedsolomon.rs:167-175 and
src/reedsolomon16.rs:282-290).  No compute calls: there is no GPU here."""
import ctypes
import re

import numpy as np
import pytest


def test_library_exports_every_declared_symbol(trs):
    from reed_solomon_amd import _lib

    declared = _lib.declared_symbols()
    assert len(declared) >= 40
    missing = [name for name in declared if not hasattr(_lib.lib, name)]
    assert not missing, missing


def test_status_codes_follow_reference_error_order(trs):
    # src/errors.rs:6-21, in declaration order, codes 1..14
    names = ["TooFewShards", "TooManyShards", "TooFewDataShards", "TooManyDataShards",
             "TooFewParityShards", "TooManyParityShards", "TooFewBufferShards",
             "TooManyBufferShards", "IncorrectShardSize", "TooFewShardsPresent", "EmptyShard",
             "InvalidShardFlags", "InvalidIndex", "SingularMatrix"]
    assert trs.ERROR_NAMES[1:] == names
    with open(trs._lib.HEADER_PATH) as f:
        header = f.read()
    codes = dict((m.group(1), int(m.group(2))) for m in re.finditer(r"TRS_([A-Z_]+) = (-?\d+)", header))
    snake = ["TOO_FEW_SHARDS", "TOO_MANY_SHARDS", "TOO_FEW_DATA_SHARDS", "TOO_MANY_DATA_SHARDS",
             "TOO_FEW_PARITY_SHARDS", "TOO_MANY_PARITY_SHARDS", "TOO_FEW_BUFFER_SHARDS",
             "TOO_MANY_BUFFER_SHARDS", "INCORRECT_SHARD_SIZE", "TOO_FEW_SHARDS_PRESENT",
             "EMPTY_SHARD", "INVALID_SHARD_FLAGS", "INVALID_INDEX", "SINGULAR_MATRIX"]
    assert [codes[s] for s in snake] == list(range(1, 15))
    assert "first shard provided is of zero length" in trs.Error(11).args[0]


@pytest.mark.parametrize("cls,k,m,err", [
    ("ReedSolomon", 0, 4, "TooFewDataShards"),
    ("ReedSolomon", 4, 0, "TooFewParityShards"),
    ("ReedSolomon", 200, 57, "TooManyShards"),
    ("ReedSolomon16", 0, 4, "TooFewDataShards"),
    ("ReedSolomon16", 4, 0, "TooFewParityShards"),
    ("ReedSolomon16", 60000, 5537, "TooManyShards"),
])
def test_new_validation_without_device(trs, cls, k, m, err):
    with pytest.raises(trs.Error) as e:
        getattr(trs, cls)(k, m)
    assert e.value.name == err


def test_slice_length_mismatch_is_rejected_before_device(trs):
    # the reference panics on unequal slices (src/gf/scalar.rs:8); the C ABI
    # returns TRS_ERR_INVALID_ARGUMENT without touching the device
    rc = trs._lib.lib.trs_gf8_mul_slice(ctypes.c_void_p(16), 10, ctypes.c_void_p(32), 11, 3, 0, None)
    assert rc == -3


def _gauss_jordan_rows(field, k, n, present):
    """The reference's plan rows by its own steps: invert the sub-generator of
    the first k present rows (src/reedsolomon.rs:537-571), missing-data rows of
    the inverse, then parity rows times the inverse (:866-889)."""
    import oracle as O

    G = O.build_matrix(field, k, n)
    valid = [i for i in range(n) if present[i]][:k]
    rc, inv = O.invert(field, np.ascontiguousarray(G[valid]))
    assert rc == 0
    data = [i for i in range(k) if not present[i]]
    par = [i for i in range(k, n) if not present[i]]
    rows = [inv[d] for d in data]
    if par:
        rows += list(O.multiply(field, np.ascontiguousarray(G[par]), inv))
    return np.array(rows).reshape(len(rows), k)


@pytest.mark.parametrize("field,k,m,seed", [
    (8, 10, 4, 1), (8, 7, 13, 2), (8, 32, 32, 3), (8, 200, 55, 4), (8, 1, 3, 5),
    (16, 4, 2, 6), (16, 17, 33, 7), (16, 200, 100, 8), (16, 300, 700, 9)])
def test_decode_matrix_equals_gauss_jordan(trs, field, k, m, seed):
    n = k + m
    rng = np.random.default_rng(seed)
    for _ in range(3):
        e = int(rng.integers(1, m + 1))
        erased = set(rng.choice(n, e, replace=False).tolist())
        present = [i not in erased for i in range(n)]
        got = trs.decode_matrix(field, k, present)
        assert (got == _gauss_jordan_rows(field, k, n, present)).all(), (field, k, m, erased)


def test_decode_matrix_c3_pattern(trs):
    # (1024,1024) with 1024 erasures: the per-pattern plan of the tower
    # workload, against the oracle's 1024 x 1024 Gauss-Jordan.
    k = m = 1024
    rng = np.random.default_rng(0xC3)
    erased = set(rng.choice(k + m, 1024, replace=False).tolist())
    present = [i not in erased for i in range(k + m)]
    got = trs.decode_matrix(16, k, present)
    assert got.shape == (1024, 1024)
    assert (got == _gauss_jordan_rows(16, k, k + m, present)).all()
