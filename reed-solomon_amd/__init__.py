"""MI355X (gfx950) backend for tape-reed-solomon's bulk coding path.

Python mirror of the reference crate's codec API over the C ABI of
``libtrs_gpu.so`` (include/trs_gpu.h).  Shards are either ``torch.uint8`` CUDA
tensors (device memory: work is enqueued on torch's current stream) or host
buffers -- uint8 numpy arrays or CPU tensors, like the reference's slices --
which the library streams through the device (the ``*_host`` entry points);
a host call returns with its outputs in host memory.  Every byte of coding
work runs in the library's HIP kernels; ``verify`` returns once its answer is
known.

Reference surface kept (tape-reed-solomon 0.2.2):

* ``ReedSolomon`` -- src/reedsolomon.rs: ``new`` (the constructor),
  ``encode``, ``encode_sep``, ``encode_rows``, ``encode_rows_into``,
  ``verify``, ``reconstruct``, ``reconstruct_data``, ``prepare_decode``,
  ``reconstruct_rows``, ``data/parity/total_shard_count``, ``encode_route``.
* ``ReedSolomon16`` -- src/reedsolomon16.rs: the same, minus
  ``encode_rows(_into)`` (absent in the reference), plus ``encode_sep_matrix``.
* ``PreparedDecoder`` / ``PreparedDecoder16`` -- ``reconstruct``,
  ``reconstruct_data``.
* ``Error`` -- src/errors.rs, same variant names and messages.
* ``mul_slice`` / ``mul_slice_xor`` -- src/gf/mod.rs:69-85.

Reconstruct accepts the reference's two shard containers: ``(tensor, bool)``
pairs (the bool is the "present" flag, src/reedsolomon.rs:44-77) and
``Optional`` entries (``None`` is missing; the codec allocates a zeroed shard
in its place, src/reedsolomon.rs:79-102).
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import lib

__all__ = [
    "Error",
    "ReedSolomon",
    "ReedSolomon16",
    "PreparedDecoder",
    "PreparedDecoder16",
    "Batch",
    "Server",
    "PackedShards",
    "pack_shards",
    "mul_slice",
    "mul_slice_xor",
    "fill_random",
    "set_host_chunk",
    "version",
]

ERROR_NAMES = [
    None,
    "TooFewShards",
    "TooManyShards",
    "TooFewDataShards",
    "TooManyDataShards",
    "TooFewParityShards",
    "TooManyParityShards",
    "TooFewBufferShards",
    "TooManyBufferShards",
    "IncorrectShardSize",
    "TooFewShardsPresent",
    "EmptyShard",
    "InvalidShardFlags",
    "InvalidIndex",
    "SingularMatrix",
]


class Error(Exception):
    """A codec error; ``.name`` is the reference's Error variant (src/errors.rs:6-21)."""

    def __init__(self, code: int):
        self.code = code
        self.name = ERROR_NAMES[code] if 0 < code < len(ERROR_NAMES) else f"Device({code})"
        super().__init__(f"{self.name}: {lib.trs_status_message(code).decode()}")

    def __eq__(self, other):
        if isinstance(other, Error):
            return self.code == other.code
        if isinstance(other, str):
            return self.name == other
        return NotImplemented

    def __hash__(self):
        return hash(self.code)


def _check(code: int) -> None:
    if code != 0:
        raise Error(code)


def version() -> str:
    return lib.trs_version().decode()


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev_ptr(t) -> int:
    import torch

    if not isinstance(t, torch.Tensor):
        raise TypeError("shards must be torch tensors")
    if t.dtype != torch.uint8 or not t.is_cuda:
        raise TypeError("shards must be torch.uint8 tensors on a CUDA (ROCm) device")
    if not t.is_contiguous():
        raise ValueError("shards must be contiguous")
    return t.data_ptr()


class PackedShards(list):
    """A list of device shards whose pointer and length arrays are built
    once: for callers that issue the same stripe many times (a server's
    clients, benchmarks), every call then skips the per-shard packing.  The
    shards must not be replaced or resized while the pack is in use."""

    def __init__(self, shards: Sequence):
        super().__init__(shards)
        self._trs_packed = _ptrs(list(self))


def pack_shards(shards: Sequence) -> PackedShards:
    return PackedShards(shards)


def _ptrs(tensors: Sequence) -> tuple:
    packed = getattr(tensors, "_trs_packed", None)
    if packed is not None:
        return packed
    arr = (ctypes.c_void_p * max(1, len(tensors)))()
    lens = (ctypes.c_size_t * max(1, len(tensors)))()
    for i, t in enumerate(tensors):
        arr[i] = _dev_ptr(t)
        lens[i] = t.numel()
    return arr, lens


def _flags(flags: Sequence[bool]):
    a = (ctypes.c_uint8 * max(1, len(flags)))()
    for i, f in enumerate(flags):
        a[i] = 1 if f else 0
    return a


def _is_host(buf) -> bool:
    """numpy arrays and CPU tensors are host shards; CUDA tensors are device
    shards."""
    import torch

    if isinstance(buf, torch.Tensor):
        return not buf.is_cuda
    if isinstance(buf, np.ndarray):
        return True
    raise TypeError("shards must be uint8 numpy arrays or torch tensors")


def _on_host(bufs) -> bool:
    """Whether a shard list lives in host memory.  Host and device buffers
    cannot be mixed in one call."""
    kinds = {_is_host(b) for b in bufs if b is not None}
    if len(kinds) > 1:
        raise TypeError("a call's shards must all be host buffers or all device tensors")
    return kinds == {True}


def _host_ptr(buf) -> tuple:
    """(address, length) of a contiguous host uint8 buffer."""
    import torch

    if isinstance(buf, torch.Tensor):
        if buf.dtype != torch.uint8 or not buf.is_contiguous():
            raise TypeError("host shards must be contiguous uint8 tensors or arrays")
        return buf.data_ptr(), buf.numel()
    if buf.dtype != np.uint8 or not buf.flags["C_CONTIGUOUS"]:
        raise TypeError("host shards must be contiguous uint8 arrays")
    if not buf.flags["WRITEABLE"]:
        raise TypeError("host shards must be writeable")
    return buf.ctypes.data, buf.size


def _hptrs(bufs: Sequence) -> tuple:
    """Pointer and length arrays of host shards (the _host entry points)."""
    arr = (ctypes.c_void_p * max(1, len(bufs)))()
    lens = (ctypes.c_size_t * max(1, len(bufs)))()
    for i, b in enumerate(bufs):
        arr[i], lens[i] = _host_ptr(b)
    return arr, lens


def _buf_ptr(buf, host: bool):
    return ctypes.c_void_p(_host_ptr(buf)[0] if host else _dev_ptr(buf))


def _numel(buf) -> int:
    return buf.size if isinstance(buf, np.ndarray) else buf.numel()


def set_host_chunk(nbytes: int = 0) -> None:
    """Bytes of each shard per pipeline chunk of host-memory calls (0 = 1 MiB)."""
    _check(lib.trs_set_host_chunk(nbytes))


def _normalise_reconstruct(shards: list, data_only: bool, k: int):
    """Turns the reference's containers into (buffers, present flags, filled
    indices, host).

    ``(buffer, bool)`` entries pass through.  ``None`` entries are missing
    Option shards: they get a zeroed buffer of the shared length, in the same
    memory as the present shards, written back into the list
    (get_or_initialize, src/reedsolomon.rs:88-101), except parity entries of a
    data-only reconstruct, which stay ``None``.
    """
    import torch

    bufs, present, fill = [], [], []
    length, device = None, None
    for i, s in enumerate(shards):
        if isinstance(s, tuple):
            t, p = s
            bufs.append(t)
            present.append(bool(p))
            if p and length is None:
                length, device = _numel(t), t
        elif s is None:
            bufs.append(None)
            present.append(False)
            fill.append(i)
        else:
            bufs.append(s)
            present.append(True)
            if length is None:
                length, device = _numel(s), s
    host = _on_host(bufs)
    if fill:
        if length is None:
            return None, present, fill, host

        def zeros():
            if host:
                return np.zeros(length, np.uint8)
            return torch.zeros(length, dtype=torch.uint8, device=device.device)

        dummy = None
        for i in fill:
            if data_only and i >= k:
                # never read or written: a data-only reconstruct skips missing parity
                if dummy is None:
                    dummy = zeros()
                bufs[i] = dummy
            else:
                bufs[i] = zeros()
    return bufs, present, fill, host


class _Codec:
    """Shards may be CUDA tensors (device memory: the work is enqueued on
    torch's current stream) or host buffers -- uint8 numpy arrays or CPU
    tensors, the reference's slices -- which stream through the device and are
    complete when the call returns.  One call takes one kind."""

    _fam = ""

    def __init__(self, data_shards: int, parity_shards: int, _handle=None):
        if _handle is not None:
            self._h = _handle
            return
        h = ctypes.c_void_p()
        _check(getattr(lib, f"trs_{self._fam}_new")(data_shards, parity_shards, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and lib is not None:
            try:
                getattr(lib, f"trs_{self._fam}_free")(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    def _fn(self, name: str, host: bool):
        return getattr(lib, f"trs_{self._fam}_{name}{'_host' if host else ''}")

    def clone(self):
        """Clone (src/reedsolomon.rs:725-739): same shape and routes, a copy of
        the plan cache."""
        h = ctypes.c_void_p()
        _check(getattr(lib, f"trs_{self._fam}_clone")(self._h, ctypes.byref(h)))
        return type(self)(0, 0, _handle=h)

    __copy__ = clone

    def data_shard_count(self) -> int:
        return getattr(lib, f"trs_{self._fam}_data_shard_count")(self._h)

    def parity_shard_count(self) -> int:
        return getattr(lib, f"trs_{self._fam}_parity_shard_count")(self._h)

    def total_shard_count(self) -> int:
        return getattr(lib, f"trs_{self._fam}_total_shard_count")(self._h)

    def __eq__(self, other):  # src/reedsolomon.rs:741-746
        return (type(self) is type(other)
                and self.data_shard_count() == other.data_shard_count()
                and self.parity_shard_count() == other.parity_shard_count())

    def encode(self, shards: Sequence) -> None:
        """Parity shards (the last m of ``shards``) are overwritten."""
        if _on_host(shards):
            p, l = _hptrs(shards)
            _check(self._fn("encode", True)(self._h, p, l, len(shards)))
            return
        p, l = _ptrs(shards)
        _check(self._fn("encode", False)(self._h, p, l, len(shards), _stream()))

    def encode_sep(self, data: Sequence, parity: Sequence) -> None:
        host = _on_host(list(data) + list(parity))
        conv = _hptrs if host else _ptrs
        dp, dl = conv(data)
        pp, pl = conv(parity)
        args = (self._h, dp, dl, len(data), pp, pl, len(parity))
        _check(self._fn("encode_sep", host)(*(args if host else args + (_stream(),))))

    def verify(self, shards: Sequence) -> bool:
        ok = ctypes.c_int(0)
        if _on_host(shards):
            p, l = _hptrs(shards)
            _check(self._fn("verify", True)(self._h, p, l, len(shards), ctypes.byref(ok)))
        else:
            p, l = _ptrs(shards)
            _check(self._fn("verify", False)(self._h, p, l, len(shards), ctypes.byref(ok),
                                             _stream()))
        return bool(ok.value)

    def _call_reconstruct(self, handle, fam: str, shards: list, data_only: bool,
                          all_missing_error: int) -> None:
        k = self.data_shard_count()
        bufs, present, filled, host = _normalise_reconstruct(shards, data_only, k)
        n = len(shards)
        if bufs is None:  # nothing present at all
            total = self.total_shard_count()
            if n != total:
                raise Error(1 if n < total else 2)
            raise Error(all_missing_error)
        p, l = (_hptrs if host else _ptrs)(bufs)
        fn = getattr(lib, f"trs_{fam}_reconstruct{'_host' if host else ''}")
        args = (handle, p, l, _flags(present), n, int(data_only))
        _check(fn(*(args if host else args + (_stream(),))))
        for i in filled:
            if not (data_only and i >= k):
                shards[i] = bufs[i]

    def _reconstruct(self, shards: list, data_only: bool) -> None:
        self._call_reconstruct(self._h, self._fam, shards, data_only, 10)

    def reconstruct(self, shards: list) -> None:
        self._reconstruct(shards, False)

    def reconstruct_data(self, shards: list) -> None:
        self._reconstruct(shards, True)

    def reconstruct_rows(self, rows, row_len: int, present: Sequence[bool]) -> None:
        host = _is_host(rows)
        args = (self._h, _buf_ptr(rows, host), _numel(rows), row_len, _flags(present),
                len(present))
        _check(self._fn("reconstruct_rows", host)(*(args if host else args + (_stream(),))))

    def prepare_decode(self, present: Sequence[bool]):
        h = ctypes.c_void_p()
        _check(getattr(lib, f"trs_{self._fam}_prepare_decode")(
            self._h, _flags(present), len(present), ctypes.byref(h)))
        return self._decoder_type(h, self)


class _Decoder:
    _fam = ""

    def __init__(self, handle, codec):
        self._h = handle
        self._codec = codec  # the plan's device memory outlives no codec
        self._k = codec.data_shard_count()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and lib is not None:
            try:
                getattr(lib, f"trs_{self._fam}_free")(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    def _reconstruct(self, shards: list, data_only: bool) -> None:
        self._codec._call_reconstruct(self._h, self._fam, shards, data_only, 12)

    def reconstruct(self, shards: list) -> None:
        self._reconstruct(shards, False)

    def reconstruct_data(self, shards: list) -> None:
        self._reconstruct(shards, True)


class PreparedDecoder(_Decoder):
    """src/reedsolomon.rs:748-797."""

    _fam = "dec8"


class PreparedDecoder16(_Decoder):
    """src/reedsolomon16.rs:709-753."""

    _fam = "dec16"


class ReedSolomon(_Codec):
    """GF(2^8) systematic Reed-Solomon, poly 0x11d (src/reedsolomon.rs)."""

    _fam = "rs8"
    _decoder_type = PreparedDecoder

    def encode_route(self, shard_len: int) -> str:
        return lib.trs_rs8_encode_route(self._h, shard_len).decode()

    def reference_route(self, shard_len: int) -> str:
        """The route the reference's x86 (GFNI) build takes for this shape and
        shard length (ReedSolomon::encode_route, src/reedsolomon.rs:239-253):
        "fft-generated", "fft-staged" or "fused-matrix"."""
        return lib.trs_rs8_reference_route(self._h, shard_len).decode()

    # "staged-fft": the GF(2^8) staged FFT program (src/fft.rs:393-592) on the
    # byte-plane LDS interpreter.
    ROUTES = {"table": 0, "bitsliced-matrix": 1, "bitsliced-fft": 2, "staged-fft": 3}

    def encode_sep_route(self, route: str, data: Sequence, parity: Sequence) -> None:
        """Encode through a named device route (the reference's encode_forced
        bench hook, src/reedsolomon.rs:365-398).  Raises Error(Device(-4)) when
        the shape has no program for that route."""
        dp, dl = _ptrs(data)
        pp, pl = _ptrs(parity)
        _check(lib.trs_rs8_encode_sep_route(self._h, self.ROUTES[route], dp, dl, len(data), pp,
                                            pl, len(parity), _stream()))

    APPLY_ROUTES = {"auto": 0, "table": 1, "mfma": 2}
    DECODE_ROUTES = ("auto", "matrix", "fft")

    def set_decode_route(self, route: str) -> None:
        """Reconstruct route: "matrix" (the reference's decode plans), "fft"
        (the additive-FFT erasure decoder over GF(2^8)) or "auto"."""
        _check(lib.trs_rs8_set_decode_route(self._h, self.DECODE_ROUTES.index(route)))

    def set_apply_route(self, route: str) -> None:
        """Device route for matrix applies (reconstruct; encode of shapes
        without a generated program): "auto", "table" or "mfma" (int8 MFMA over
        the bit-expanded rows, k <= 32).  "auto" takes MFMA for single calls
        with k <= 32 and shards under 512 KiB.  All byte-identical."""
        _check(lib.trs_rs8_set_apply_route(self._h, self.APPLY_ROUTES[route]))

    def encode_fused(self, shards: Sequence) -> None:
        """Alias of encode (src/reedsolomon.rs:354-359)."""
        self.encode(shards)

    def reconstruct_alloc(self, shards: Sequence, data_only: bool = False) -> list:
        """reconstruct over Option<T> shards (src/reedsolomon.rs:88-101): None
        entries are missing shards that the codec allocates on the device and
        rebuilds; they come back as DeviceBuffer objects (missing parity stays
        None with data_only)."""
        n = len(shards)
        arr = (ctypes.c_void_p * max(1, n))()
        lens = (ctypes.c_size_t * max(1, n))()
        for i, t in enumerate(shards):
            if t is not None:
                arr[i] = _dev_ptr(t)
                lens[i] = t.numel()
        _check(lib.trs_rs8_reconstruct_alloc(self._h, arr, lens, n, int(data_only), _stream()))
        out = []
        for i, t in enumerate(shards):
            if t is None and arr[i]:
                out.append(DeviceBuffer(arr[i], lens[i]))
            else:
                out.append(t)
        return out

    def encode_rows_into(self, data, row_len: int, parity) -> None:
        host = _on_host([data, parity])
        args = (self._h, _buf_ptr(data, host), _numel(data), row_len, _buf_ptr(parity, host),
                _numel(parity))
        _check(self._fn("encode_rows_into", host)(*(args if host else args + (_stream(),))))

    def encode_rows(self, data, row_len: int):
        """Panics (AssertionError) on a bad length, like src/reedsolomon.rs:652-657.
        Returns the parity rows in the same memory as ``data``."""
        import torch

        assert _numel(data) == self.data_shard_count() * row_len, \
            "data must be data_shard_count * row_len bytes"
        n = self.parity_shard_count() * row_len
        if _is_host(data):
            parity = np.zeros(n, np.uint8)
        else:
            parity = torch.zeros(n, dtype=torch.uint8, device=data.device)
        self.encode_rows_into(data, row_len, parity)
        return parity

    def generator(self) -> np.ndarray:
        n, k = self.total_shard_count(), self.data_shard_count()
        out = np.zeros(n * k, np.uint8)
        _check(lib.trs_rs8_generator(self._h, out.ctypes.data_as(_lib.c_u8p), out.size))
        return out.reshape(n, k)


class ReedSolomon16(_Codec):
    """GF((2^8)^2) tower Reed-Solomon over split-plane shards (src/reedsolomon16.rs)."""

    _fam = "rs16"
    _decoder_type = PreparedDecoder16

    def encode_route(self) -> str:
        return lib.trs_rs16_encode_route(self._h).decode()

    def encode_sep_matrix(self, data: Sequence, parity: Sequence) -> None:
        dp, dl = _ptrs(data)
        pp, pl = _ptrs(parity)
        _check(lib.trs_rs16_encode_sep_matrix(self._h, dp, dl, len(data), pp, pl, len(parity),
                                              _stream()))

    def encode_sep_fft(self, data: Sequence, parity: Sequence) -> None:
        """The tower-FFT route regardless of routing (src/fft16.rs:887-930)."""
        dp, dl = _ptrs(data)
        pp, pl = _ptrs(parity)
        _check(lib.trs_rs16_encode_sep_fft(self._h, dp, dl, len(data), pp, pl, len(parity),
                                           _stream()))

    ENCODE_ROUTES = ("auto", "matrix", "fft")
    DECODE_ROUTES = ("auto", "matrix", "fft")

    def set_decode_route(self, route: str) -> None:
        """Reconstruct route: "matrix" (the reference's decode plans), "fft"
        (additive-FFT erasure decoder) or "auto"."""
        _check(lib.trs_rs16_set_decode_route(self._h, self.DECODE_ROUTES.index(route)))

    APPLY_ROUTES = {"auto": 0, "table": 1, "mfma": 2}

    def set_apply_route(self, route: str) -> None:
        """Device route for tower matrix applies (matrix-route encode,
        encode_sep_matrix, matrix-plan reconstruct): "auto", "table" (the
        Karatsuba kernel) or "mfma" (int8 MFMA over the 16m x 16k bit
        expansion; k <= 16).  All byte-identical."""
        _check(lib.trs_rs16_set_apply_route(self._h, self.APPLY_ROUTES[route]))

    def reconstruct_stripes(self, stripes, present, data_only: bool = False) -> None:
        """Rebuilds the missing shards of many stripes in one call: ``stripes``
        is a uint8 device tensor (stripes, k + m, length) with a contiguous last
        dimension, ``present`` a (stripes, k + m) boolean array."""
        t = stripes
        n = self.total_shard_count()
        if t.dim() != 3 or t.shape[1] != n or t.stride(2) != 1 or t.dtype.itemsize != 1:
            raise Error(-3)
        flags = np.ascontiguousarray(np.asarray(present, dtype=bool), dtype=np.uint8)
        if flags.shape != (t.shape[0], n):
            raise Error(-3)
        idx = np.arange(t.shape[0], dtype=np.uint64)[:, None] * np.uint64(t.stride(0)) + \
            np.arange(n, dtype=np.uint64)[None, :] * np.uint64(t.stride(1))
        ptrs = (idx + np.uint64(t.data_ptr())).reshape(-1)
        _check(lib.trs_rs16_reconstruct_stripes(
            self._h, ptrs.ctypes.data_as(ctypes.POINTER(ctypes.c_void_p)),
            flags.ctypes.data_as(_lib.c_u8p), t.shape[0], t.shape[2], int(data_only), _stream()))

    def encode_stripes(self, stripes, route: str = "auto") -> None:
        """Encodes many stripes (each k data then m parity tensors of one
        common even length) in one call; ``route`` as ENCODE_ROUTES.
        ``stripes`` is a list of shard lists, or one uint8 device tensor of
        shape (stripes, k + m, length) whose last dimension is contiguous."""
        n = self.total_shard_count()
        if hasattr(stripes, "dim"):
            t = stripes
            if t.dim() != 3 or t.shape[1] != n or t.stride(2) != 1 or t.dtype.itemsize != 1:
                raise Error(-3)
            idx = np.arange(t.shape[0], dtype=np.uint64)[:, None] * np.uint64(t.stride(0)) + \
                np.arange(n, dtype=np.uint64)[None, :] * np.uint64(t.stride(1))
            ptrs = (idx + np.uint64(t.data_ptr())).reshape(-1)
            arr = ptrs.ctypes.data_as(ctypes.POINTER(ctypes.c_void_p))
            _check(lib.trs_rs16_encode_stripes(self._h, arr, t.shape[0], t.shape[2],
                                               self.ENCODE_ROUTES.index(route), _stream()))
            return
        flat = [t for s in stripes for t in s]
        if any(len(s) != n for s in stripes):
            raise Error(-3)
        length = flat[0].numel() if flat else 0
        if any(t.numel() != length for t in flat):
            raise Error(ERROR_NAMES.index("IncorrectShardSize"))
        p, _ = _ptrs(flat)
        _check(lib.trs_rs16_encode_stripes(self._h, p, len(stripes), length,
                                           self.ENCODE_ROUTES.index(route), _stream()))

    def generator(self) -> np.ndarray:
        n, k = self.total_shard_count(), self.data_shard_count()
        out = np.zeros(n * k, np.uint16)
        _check(lib.trs_rs16_generator(self._h, out.ctypes.data_as(_lib.c_u16p), out.size))
        return out.reshape(n, k)


def decode_matrix(field: int, k: int, present: Sequence[bool]) -> np.ndarray:
    """The fused decode rows a reconstruct with these presence flags applies
    (missing data rows, then missing parity rows, over the first k present
    shards); computed on the host."""
    n = len(present)
    rows = ctypes.c_size_t(0)
    flags = _flags(present)
    _check(lib.trs_decode_matrix(field, k, n, flags, None, 0, ctypes.byref(rows)))
    out = np.zeros((rows.value, k), np.uint8 if field == 8 else np.uint16)
    _check(lib.trs_decode_matrix(field, k, n, flags, out.ctypes.data_as(ctypes.c_void_p), out.size,
                                 ctypes.byref(rows)))
    return out


def fft16_program_info(k: int, m: int) -> Optional[dict]:
    """Host-side statistics of the staged tower-FFT program for (k, m) and of
    its device lowering; None when the shape has no program."""
    out = (ctypes.c_size_t * 9)()
    rc = lib.trs_fft16_program_info(k, m, out, 9)
    if rc == -4:
        return None
    _check(rc)
    keys = ("register_count", "mult_count", "glue", "stages", "zero_regs", "profitable",
            "device_regs", "device_steps", "group_bytes")
    return {key: int(v) for key, v in zip(keys, out)}


def fft8_program_info(k: int, m: int) -> Optional[dict]:
    """Host-side statistics of the GF(2^8) staged FFT program for (k, m)
    (src/fft.rs:393-592) and of its device lowering; `reference_route` is 1
    when the reference's x86 build routes the shape to it
    (src/reedsolomon.rs:204-234).  None outside 1 <= k, 1 <= m, k + m <= 256."""
    out = (ctypes.c_size_t * 10)()
    rc = lib.trs_fft8_program_info(k, m, out, 10)
    if rc == -4:
        return None
    _check(rc)
    keys = ("register_count", "mult_count", "glue", "stages", "zero_regs", "reference_route",
            "device_regs", "device_steps", "group_bytes", "device_mult_count")
    return {key: int(v) for key, v in zip(keys, out)}


class Batch:
    """Many GF(2^8) encode / reconstruct requests run as one kernel per work list.

    No reference counterpart: the reference codes one stripe per call and leaves
    batching to callers (BENCH-RESULTS.md:98-100).  Each queued request is
    validated like the single call and produces byte-identical output.
    """

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib.trs_batch8_new(ctypes.byref(h)))
        self._h = h
        self._keep = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and lib is not None:
            try:
                lib.trs_batch8_free(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    def clear(self) -> None:
        lib.trs_batch8_clear(self._h)
        self._keep = []

    def add_encode(self, rs: ReedSolomon, shards: Sequence) -> None:
        p, l = _ptrs(shards)
        _check(lib.trs_batch8_add_encode(self._h, rs._h, p, l, len(shards)))
        self._keep.append(rs)

    def add_reconstruct(self, rs: ReedSolomon, shards: Sequence, present: Sequence[bool],
                        data_only: bool = False) -> None:
        p, l = _ptrs(shards)
        _check(lib.trs_batch8_add_reconstruct(self._h, rs._h, p, l, _flags(present), len(shards),
                                              int(data_only)))
        self._keep.append(rs)

    def prepare(self) -> None:
        _check(lib.trs_batch8_prepare(self._h))

    def set_graph(self, enable: bool = True) -> None:
        """Replay the prepared launches as one captured hipGraph."""
        _check(lib.trs_batch8_set_graph(self._h, int(enable)))

    def run(self, stream=None) -> None:
        s = _stream() if stream is None else ctypes.c_void_p(stream)
        _check(lib.trs_batch8_run(self._h, s))

    def bytes(self) -> int:
        return lib.trs_batch8_bytes(self._h)

    def __len__(self) -> int:
        return lib.trs_batch8_size(self._h)


class Server:
    """Continuous-batching request loop over device-resident shards (no
    reference counterpart).  submit_* validate like the single calls and return
    a ticket; wait(ticket) blocks until that request is done; take(ticket)
    returns (latency_us, bytes).  Tensors must stay alive until completion."""

    def __init__(self, slots: int = 4, hbm_budget_bytes: int = 64 << 30, max_batch: int = 256,
                 window_us: int = 50):
        h = ctypes.c_void_p()
        _check(lib.trs_server_new(slots, hbm_budget_bytes, max_batch, window_us,
                                  ctypes.byref(h)))
        self._h = h
        self._keep = {}

    def close(self) -> None:
        h = getattr(self, "_h", None)
        if h and lib is not None:
            try:
                lib.trs_server_free(h)
            except Exception:
                pass
            self._h = None
            self._keep = {}

    def __del__(self):
        self.close()

    def _ticket(self, rc: int, t, keep) -> int:
        _check(rc)
        self._keep[t.value] = keep
        return t.value

    def submit_encode(self, rs, shards: Sequence) -> int:
        p, l = _ptrs(shards)
        t = ctypes.c_uint64()
        fn = lib.trs_server_submit_encode16 if isinstance(rs, ReedSolomon16) else \
            lib.trs_server_submit_encode8
        return self._ticket(fn(self._h, rs._h, p, l, len(shards), ctypes.byref(t)), t,
                            (rs, shards))

    def submit_reconstruct(self, rs, shards: Sequence, present: Sequence[bool],
                           data_only: bool = False) -> int:
        p, l = _ptrs(shards)
        t = ctypes.c_uint64()
        fn = lib.trs_server_submit_reconstruct16 if isinstance(rs, ReedSolomon16) else \
            lib.trs_server_submit_reconstruct8
        return self._ticket(fn(self._h, rs._h, p, l, _flags(present), len(shards), int(data_only),
                               ctypes.byref(t)), t, (rs, shards))

    def wait(self, ticket: int) -> None:
        st = ctypes.c_int(0)
        _check(lib.trs_server_wait(self._h, ticket, ctypes.byref(st)))
        _check(st.value)

    def poll(self, ticket: int) -> bool:
        done, st = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib.trs_server_poll(self._h, ticket, ctypes.byref(done), ctypes.byref(st)))
        if done.value:
            _check(st.value)
        return bool(done.value)

    def drain(self) -> None:
        _check(lib.trs_server_drain(self._h))

    def take(self, ticket: int):
        lat, nb, st = ctypes.c_double(0), ctypes.c_uint64(0), ctypes.c_int(0)
        _check(lib.trs_server_take(self._h, ticket, ctypes.byref(lat), ctypes.byref(nb),
                                   ctypes.byref(st)))
        self._keep.pop(ticket, None)
        _check(st.value)
        return lat.value, nb.value

    def take_detail(self, ticket: int):
        """(latency_us, (queued_us, launch_us, device_us), bytes) of a finished
        ticket: queued = submit to taken into a batch, launch = host-side build
        and enqueue, device = enqueued to seen complete."""
        lat, nb, st = ctypes.c_double(0), ctypes.c_uint64(0), ctypes.c_int(0)
        ph = (ctypes.c_double * 3)()
        _check(lib.trs_server_take_detail(self._h, ticket, ctypes.byref(lat), ph,
                                          ctypes.byref(nb), ctypes.byref(st)))
        self._keep.pop(ticket, None)
        _check(st.value)
        return lat.value, (ph[0], ph[1], ph[2]), nb.value

    def set_hbm_reserve(self, reserve_bytes: int) -> None:
        """Free HBM the server's own batch tables must leave untouched."""
        _check(lib.trs_server_set_hbm_reserve(self._h, reserve_bytes))

    def stats(self) -> dict:
        """Counters: requests completed, batches launched, batches split for
        memory, requests failed for memory."""
        out = (ctypes.c_uint64 * 4)()
        _check(lib.trs_server_stats(self._h, out))
        return {"done": out[0], "batches": out[1], "splits": out[2], "oom": out[3]}

    def set_spin(self, spin_us: int) -> None:
        """Microseconds the server's threads poll before blocking (0 = block)."""
        _check(lib.trs_server_set_spin(self._h, spin_us))


def _slice(out, inp, c: int, xor: bool) -> None:
    _check(lib.trs_gf8_mul_slice(ctypes.c_void_p(_dev_ptr(out)), out.numel(),
                                 ctypes.c_void_p(_dev_ptr(inp)), inp.numel(), c, int(xor),
                                 _stream()))


def mul_slice(out, inp, c: int) -> None:
    """out = c * inp (src/gf/mod.rs:71-76).  Unequal lengths raise like the reference's panic."""
    if out.numel() != inp.numel():
        raise AssertionError("assertion `left == right` failed: slice lengths differ")
    _slice(out, inp, c, False)


def mul_slice_xor(out, inp, c: int) -> None:
    """out ^= c * inp (src/gf/mod.rs:80-85)."""
    if out.numel() != inp.numel():
        raise AssertionError("assertion `left == right` failed: slice lengths differ")
    _slice(out, inp, c, True)


class DeviceBuffer:
    """A shard buffer the library allocated on the device (trs_rs8_reconstruct_alloc)."""

    def __init__(self, ptr: int, nbytes: int):
        self.ptr, self.nbytes = int(ptr), int(nbytes)

    def to_tensor(self):
        """A torch copy of the buffer (through the slice kernel with c = 1)."""
        import torch

        t = torch.empty(self.nbytes, dtype=torch.uint8, device="cuda")
        _check(lib.trs_gf8_mul_slice(ctypes.c_void_p(t.data_ptr()), self.nbytes,
                                     ctypes.c_void_p(self.ptr), self.nbytes, 1, 0, _stream()))
        return t

    def free(self) -> None:
        if self.ptr:
            import torch

            torch.cuda.synchronize()
            _check(lib.trs_device_free(ctypes.c_void_p(self.ptr)))
            self.ptr = 0


def u16_to_split(symbols, shard) -> None:
    """Interleaved little-endian u16 symbols (a uint8 tensor of 2n bytes) to the
    split-plane layout [low n | high n] (src/reedsolomon16.rs:4-7)."""
    if symbols.numel() != shard.numel() or symbols.numel() % 2:
        raise ValueError("both buffers hold 2 * n_symbols bytes")
    _check(lib.trs_u16_to_split(ctypes.c_void_p(_dev_ptr(symbols)), symbols.numel() // 2,
                                ctypes.c_void_p(_dev_ptr(shard)), _stream()))


def split_to_u16(shard, symbols) -> None:
    """The split-plane layout back to interleaved little-endian u16 symbols."""
    if symbols.numel() != shard.numel() or shard.numel() % 2:
        raise ValueError("both buffers hold 2 * n_symbols bytes")
    _check(lib.trs_split_to_u16(ctypes.c_void_p(_dev_ptr(shard)), shard.numel() // 2,
                                ctypes.c_void_p(_dev_ptr(symbols)), _stream()))


def fill_random(t, seed: int) -> None:
    """Fill a device tensor with the seeded splitmix64 byte stream."""
    _check(lib.trs_fill_random(ctypes.c_void_p(_dev_ptr(t)), t.numel(),
                               seed & 0xFFFFFFFFFFFFFFFF, _stream()))
