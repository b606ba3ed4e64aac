"""ctypes binding of libtrs_gpu.so (the C ABI declared in include/trs_gpu.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C reed-solomon_amd/csrc``).  There is no fallback: if the shared
library is missing, importing the package fails with an error that says so.
"""
from __future__ import annotations

import ctypes
import os

# torch ships its own libamdhip64 (same SONAME, libamdhip64.so.7).  Loading
# torch first makes libtrs_gpu.so bind to that already-loaded HIP runtime;
# loading ours first would put two HIP runtimes in the process and torch would
# then see no GPU.
import torch  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
# TRS_GPU_LIB names another build of the library (for comparing builds).
LIB_PATH = os.environ.get("TRS_GPU_LIB") or os.path.join(HERE, "libtrs_gpu.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "trs_gpu.h")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
        "or `make -C reed-solomon_amd/csrc`"
    )

lib = ctypes.CDLL(LIB_PATH)

c_size = ctypes.c_size_t
c_int = ctypes.c_int
c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_u16p = ctypes.POINTER(ctypes.c_uint16)
c_vp = ctypes.c_void_p
c_sizep = ctypes.POINTER(ctypes.c_size_t)
c_ptrs = ctypes.POINTER(ctypes.c_void_p)


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("trs_status_message", ctypes.c_char_p, c_int)
_sig("trs_version", ctypes.c_char_p)
_sig("trs_gf8_mul_slice", c_int, c_vp, c_size, c_vp, c_size, ctypes.c_uint8, c_int, c_vp)
_sig("trs_fill_random", c_int, c_vp, c_size, ctypes.c_uint64, c_vp)
_sig("trs_rs8_reconstruct_alloc", c_int, c_vp, c_ptrs, c_sizep, c_size, c_int, c_vp)
_sig("trs_device_alloc", c_int, c_size, ctypes.POINTER(c_vp))
_sig("trs_device_free", c_int, c_vp)
_sig("trs_memcpy", c_int, c_vp, c_vp, c_size, c_vp)
_sig("trs_stream_synchronize", c_int, c_vp)
_sig("trs_set_host_chunk", c_int, c_size)
_sig("trs_u16_to_split", c_int, c_vp, c_size, c_vp, c_vp)
_sig("trs_split_to_u16", c_int, c_vp, c_size, c_vp, c_vp)

for fam, elem in (("rs8", c_u8p), ("rs16", c_u16p)):
    _sig(f"trs_{fam}_new", c_int, c_size, c_size, ctypes.POINTER(c_vp))
    _sig(f"trs_{fam}_free", None, c_vp)
    for what in ("data", "parity", "total"):
        _sig(f"trs_{fam}_{what}_shard_count", c_size, c_vp)
    _sig(f"trs_{fam}_generator", c_int, c_vp, elem, c_size)
    _sig(f"trs_{fam}_encode", c_int, c_vp, c_ptrs, c_sizep, c_size, c_vp)
    _sig(f"trs_{fam}_encode_sep", c_int, c_vp, c_ptrs, c_sizep, c_size, c_ptrs, c_sizep, c_size,
         c_vp)
    _sig(f"trs_{fam}_verify", c_int, c_vp, c_ptrs, c_sizep, c_size, ctypes.POINTER(c_int), c_vp)
    _sig(f"trs_{fam}_reconstruct", c_int, c_vp, c_ptrs, c_sizep, c_u8p, c_size, c_int, c_vp)
    _sig(f"trs_{fam}_reconstruct_rows", c_int, c_vp, c_vp, c_size, c_size, c_u8p, c_size, c_vp)
    _sig(f"trs_{fam}_prepare_decode", c_int, c_vp, c_u8p, c_size, ctypes.POINTER(c_vp))
    _sig(f"trs_{fam}_clone", c_int, c_vp, ctypes.POINTER(c_vp))
    # host-memory forms: the same arguments without the stream
    _sig(f"trs_{fam}_encode_host", c_int, c_vp, c_ptrs, c_sizep, c_size)
    _sig(f"trs_{fam}_encode_sep_host", c_int, c_vp, c_ptrs, c_sizep, c_size, c_ptrs, c_sizep,
         c_size)
    _sig(f"trs_{fam}_verify_host", c_int, c_vp, c_ptrs, c_sizep, c_size, ctypes.POINTER(c_int))
    _sig(f"trs_{fam}_reconstruct_host", c_int, c_vp, c_ptrs, c_sizep, c_u8p, c_size, c_int)
    _sig(f"trs_{fam}_reconstruct_rows_host", c_int, c_vp, c_vp, c_size, c_size, c_u8p, c_size)

_sig("trs_rs8_encode_route", ctypes.c_char_p, c_vp, c_size)
_sig("trs_rs8_reference_route", ctypes.c_char_p, c_vp, c_size)
_sig("trs_rs8_set_decode_route", c_int, c_vp, c_int)
_sig("trs_rs16_encode_route", ctypes.c_char_p, c_vp)
_sig("trs_rs8_encode_sep_route", c_int, c_vp, c_int, c_ptrs, c_sizep, c_size, c_ptrs, c_sizep,
     c_size, c_vp)
_sig("trs_rs8_set_apply_route", c_int, c_vp, c_int)
_sig("trs_rs8_encode_rows_into", c_int, c_vp, c_vp, c_size, c_size, c_vp, c_size, c_vp)
_sig("trs_rs8_encode_rows_into_host", c_int, c_vp, c_vp, c_size, c_size, c_vp, c_size)
_sig("trs_rs16_encode_sep_matrix", c_int, c_vp, c_ptrs, c_sizep, c_size, c_ptrs, c_sizep, c_size,
     c_vp)
_sig("trs_rs16_encode_sep_fft", c_int, c_vp, c_ptrs, c_sizep, c_size, c_ptrs, c_sizep, c_size,
     c_vp)
_sig("trs_rs16_encode_stripes", c_int, c_vp, c_ptrs, c_size, c_size, c_int, c_vp)
_sig("trs_rs16_encode_stripes_noalloc", c_int, c_vp, c_ptrs, c_size, c_size, c_vp)
_sig("trs_rs16_set_decode_route", c_int, c_vp, c_int)
_sig("trs_rs16_set_apply_route", c_int, c_vp, c_int)
_sig("trs_rs16_reconstruct_stripes", c_int, c_vp, c_ptrs, c_u8p, c_size, c_size, c_int, c_vp)
_sig("trs_fft16_program_info", c_int, c_size, c_size, ctypes.POINTER(c_size), c_size)
_sig("trs_decode_matrix", c_int, c_int, c_size, c_size, c_u8p, c_vp, c_size,
     ctypes.POINTER(c_size))
c_u32p = ctypes.POINTER(ctypes.c_uint32)
_sig("trs_fft16_lowered_program", c_int, c_size, c_size, c_u32p, c_size, c_u32p, c_size, c_u32p,
     c_size, c_u32p, c_size, ctypes.POINTER(c_size))
_sig("trs_fft8_program_info", c_int, c_size, c_size, ctypes.POINTER(c_size), c_size)
_sig("trs_fft8_lowered_program", c_int, c_size, c_size, c_u32p, c_size, c_u32p, c_size, c_u32p,
     c_size, c_u32p, c_size, ctypes.POINTER(c_size))
_sig("trs_fft8_decode_program", c_int, c_size, c_size, c_u8p, c_u32p, c_size, c_u32p, c_size,
     c_u32p, c_size, c_u32p, c_size, c_u32p, c_size, ctypes.POINTER(c_size))
_sig("trs_fft16_decode_program", c_int, c_size, c_size, c_u8p, c_u32p, c_size, c_u32p, c_size,
     c_u32p, c_size, c_u32p, c_size, c_u32p, c_size, ctypes.POINTER(c_size))
for fam in ("dec8", "dec16"):
    _sig(f"trs_{fam}_reconstruct", c_int, c_vp, c_ptrs, c_sizep, c_u8p, c_size, c_int, c_vp)
    _sig(f"trs_{fam}_reconstruct_host", c_int, c_vp, c_ptrs, c_sizep, c_u8p, c_size, c_int)
    _sig(f"trs_{fam}_free", None, c_vp)

_sig("trs_batch8_new", c_int, ctypes.POINTER(c_vp))
_sig("trs_batch8_free", None, c_vp)
_sig("trs_batch8_clear", None, c_vp)
_sig("trs_batch8_add_encode", c_int, c_vp, c_vp, c_ptrs, c_sizep, c_size)
_sig("trs_batch8_add_reconstruct", c_int, c_vp, c_vp, c_ptrs, c_sizep, c_u8p, c_size, c_int)
_sig("trs_batch8_prepare", c_int, c_vp)
_sig("trs_batch8_run", c_int, c_vp, c_vp)
_sig("trs_batch8_set_graph", c_int, c_vp, c_int)
_sig("trs_batch8_bytes", ctypes.c_uint64, c_vp)
_sig("trs_batch8_size", c_size, c_vp)


c_fp = ctypes.POINTER(ctypes.c_float)
_sig("trs_tune_apply", c_int, c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, c_int,
     c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double))
_sig("trs_tune_feasible", c_int, c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
     ctypes.POINTER(ctypes.c_uint32))
_sig("trs_tune_rule", c_int, c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
     ctypes.POINTER(c_int))
_sig("trs_tune_features", c_int, c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, c_int,
     c_fp, c_size)
_sig("trs_tuner_install", c_int, c_fp, c_size, c_fp, c_size, c_int)
_sig("trs_tuner_remove", c_int)
_sig("trs_tuner_lookup", c_int, c_int, c_int, c_size, c_size, c_size, ctypes.POINTER(c_int))
_sig("trs_mlp_param_count", c_int)
_sig("trs_mlp_new", c_int, ctypes.c_uint64, ctypes.POINTER(c_vp))
_sig("trs_mlp_free", None, c_vp)
_sig("trs_mlp_get_params", c_int, c_vp, c_fp, c_size)
_sig("trs_mlp_set_params", c_int, c_vp, c_fp, c_size)
_sig("trs_mlp_gradients", c_int, c_vp, c_fp, c_fp, c_size, c_fp, c_fp)
_sig("trs_mlp_train", c_int, c_vp, c_fp, c_fp, c_size, c_int, ctypes.c_float, c_fp)
_sig("trs_mlp_predict", c_int, c_vp, c_fp, c_size, c_fp)


c_u64p = ctypes.POINTER(ctypes.c_uint64)
_sig("trs_server_new", c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
     ctypes.POINTER(c_vp))
_sig("trs_server_free", None, c_vp)
_sig("trs_server_submit_encode8", c_int, c_vp, c_vp, c_ptrs, c_sizep, c_size, c_u64p)
_sig("trs_server_submit_reconstruct8", c_int, c_vp, c_vp, c_ptrs, c_sizep, c_u8p, c_size, c_int,
     c_u64p)
_sig("trs_server_submit_encode16", c_int, c_vp, c_vp, c_ptrs, c_sizep, c_size, c_u64p)
_sig("trs_server_submit_reconstruct16", c_int, c_vp, c_vp, c_ptrs, c_sizep, c_u8p, c_size, c_int,
     c_u64p)
_sig("trs_server_poll", c_int, c_vp, ctypes.c_uint64, ctypes.POINTER(c_int), ctypes.POINTER(c_int))
_sig("trs_server_wait", c_int, c_vp, ctypes.c_uint64, ctypes.POINTER(c_int))
_sig("trs_server_drain", c_int, c_vp)
_sig("trs_server_take", c_int, c_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double), c_u64p,
     ctypes.POINTER(c_int))
_sig("trs_server_take_detail", c_int, c_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
     ctypes.POINTER(ctypes.c_double), c_u64p, ctypes.POINTER(c_int))
_sig("trs_server_set_hbm_reserve", c_int, c_vp, ctypes.c_uint64)
_sig("trs_server_set_spin", c_int, c_vp, ctypes.c_uint32)
_sig("trs_server_stats", c_int, c_vp, c_u64p)


def declared_symbols() -> list:
    """Every function name include/trs_gpu.h declares."""
    import re

    with open(HEADER_PATH, encoding="utf-8") as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(trs_[a-z0-9_]+)\s*\(", text)))
