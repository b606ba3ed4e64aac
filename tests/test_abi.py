"""CPU checks of the drop-in boundary: the C-ABI library loads, exports every
symbol include/trs_gpu.h declares, and validates arguments that need no device
(ReedSolomon::new errors, src/reedsolomon.rs:167-175 and
src/reedsolomon16.rs:282-290).  No compute calls: there is no GPU here."""
import ctypes
import re

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
