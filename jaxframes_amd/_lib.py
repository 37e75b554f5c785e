"""ctypes binding of libjfx.so (the HIP kernels behind include/jfx.h).

The library is built in-tree by ``jaxframes_amd.build``. There is no CPU
fallback: if the library or a GPU is missing, every op raises.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libjfx.so"

JFX_U32, JFX_I32, JFX_F32, JFX_U64, JFX_I64, JFX_F64 = range(6)
JFX_ERR_CAPACITY = 6
AGG_SUM, AGG_COUNT, AGG_MIN, AGG_MAX, AGG_MEAN = 1, 2, 4, 8, 16
CMP_CODES = {">": 0, ">=": 1, "<": 2, "<=": 3, "==": 4, "!=": 5}
# device error word bits (csrc/common.hpp)
DEVERR_LOOKBACK = 1   # a bounded cross-workgroup wait ran out
DEVERR_BOUNDS = 2     # a computed position or key fell outside its declared range

_ERRNAMES = {1: "invalid argument", 2: "HIP error", 3: "unsupported", 4: "workspace too small",
             5: "device error", 6: "output capacity too small"}


class JfxError(RuntimeError):
    """A libjfx call failed."""


JOIN_MAXCOLS = 8


class JoinSide(ctypes.Structure):
    """jfx_join_side (include/jfx.h)."""
    _fields_ = [("rows", ctypes.c_void_p), ("out_rows", ctypes.c_void_p),
                ("ncols", ctypes.c_int32), ("by_row", ctypes.c_int32),
                ("record_bytes", ctypes.c_int32), ("records", ctypes.c_void_p),
                ("cols", ctypes.c_void_p * JOIN_MAXCOLS),
                ("col_bytes", ctypes.c_int32 * JOIN_MAXCOLS),
                ("col_offset", ctypes.c_int32 * JOIN_MAXCOLS),
                ("out_cols", ctypes.c_void_p * JOIN_MAXCOLS)]


BATCH_NARROW_SPAN = 1024
BATCH_CHUNK_ROWS = 65536


class BatchedRouting(ctypes.Structure):
    """jfx_batched_routing (include/jfx.h): device arrays of a routed plan."""
    _fields_ = [("n_narrow", ctypes.c_uint32), ("n_chunks", ctypes.c_uint32),
                ("n_wide", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("n_wide_rows", ctypes.c_uint64), ("part_entries", ctypes.c_uint64),
                ("merged_entries", ctypes.c_uint64),
                ("narrow_frame", ctypes.c_void_p), ("narrow_omin", ctypes.c_void_p),
                ("narrow_span", ctypes.c_void_p), ("narrow_chunk0", ctypes.c_void_p),
                ("narrow_mbase", ctypes.c_void_p), ("chunk_row0", ctypes.c_void_p),
                ("chunk_rows", ctypes.c_void_p), ("chunk_narrow", ctypes.c_void_p),
                ("chunk_part", ctypes.c_void_p), ("wide_frame", ctypes.c_void_p),
                ("wide_offsets", ctypes.c_void_p), ("frame_narrow", ctypes.c_void_p)]


_lib = None
_lock = threading.Lock()

c_u64 = ctypes.c_uint64
c_int = ctypes.c_int
c_vp = ctypes.c_void_p
c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)

# name -> argtypes (every function returns int status)
_SIGNATURES = {
    "jfx_version": [],
    "jfx_device_error": [c_vp, c_u32p],
    "jfx_sort_workspace_size": [c_int, c_int, c_u64, c_u64p],
    "jfx_sort": [c_int, c_vp, c_vp, c_int, c_int, c_vp, c_vp, c_u64, c_int, c_int, c_vp, c_u64,
                 c_vp],
    "jfx_segreduce_workspace_size": [c_int, c_u64, c_u64p],
    "jfx_segreduce": [c_int, c_vp, c_int, c_vp, c_vp, c_u64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                      c_u64p, c_vp, c_u64, c_vp],
    "jfx_prof_enable": [c_int],
    "jfx_prof_reset": [],
    "jfx_prof_filter": [ctypes.c_char_p],
    "jfx_prof_sample": [ctypes.c_uint32],
    "jfx_prof_count": [ctypes.POINTER(c_int)],
    "jfx_prof_get": [c_int, ctypes.c_char_p, c_u64, ctypes.POINTER(ctypes.c_double), c_u64p],
    "jfx_hash_rows": [ctypes.POINTER(c_vp), ctypes.POINTER(c_int), c_int, c_u64, c_vp, c_vp],
    "jfx_hash_partition_workspace_size": [c_int, c_u64, ctypes.c_uint32, c_u64p],
    "jfx_hash_partition": [c_int, c_vp, c_u64, ctypes.c_uint32, c_int, c_vp, c_vp, c_vp, c_vp,
                           c_vp, c_u64, c_vp],
    "jfx_join_workspace_size": [c_u64, c_u64p],
    "jfx_join_count": [c_int, c_vp, c_u64, c_vp, c_u64, c_u64p, c_vp, c_u64, c_vp],
    "jfx_join_fill": [c_int, c_vp, c_vp, c_u64, c_vp, c_vp, c_u64, c_vp, c_vp, c_vp, c_vp, c_u64,
                      c_vp],
    "jfx_join_fill_values": [c_int, c_vp, c_vp, c_u64, c_vp, c_vp, c_u64, c_int, c_vp, c_int, c_int,
                             c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_vp],
    "jfx_join_fill_columns": [c_int, c_vp, c_u64, c_vp, c_vp, c_u64, c_vp, c_vp, c_vp, c_u64, c_vp],
    "jfx_pack_columns": [c_int, ctypes.POINTER(c_vp), ctypes.POINTER(c_int), c_u64, c_vp,
                         ctypes.POINTER(c_int), c_vp],
    "jfx_join_prefilter_workspace_size": [c_u64, c_u64p],
    "jfx_join_prefilter": [c_int, c_vp, c_u64, c_vp, c_u64, c_int, ctypes.POINTER(c_vp),
                           ctypes.POINTER(c_int), c_vp, c_vp, c_vp, ctypes.POINTER(c_u64), c_vp,
                           c_u64, c_vp],
    "jfx_join_first_match": [c_int, c_vp, c_u64, c_vp, c_vp, c_u64, c_vp, c_vp],
    "jfx_gather": [c_int, c_vp, c_vp, c_u64, c_vp, c_u64, c_vp],
    "jfx_batched_groupby_workspace_size": [ctypes.c_uint32, c_int, c_u64, c_u64p],
    "jfx_batched_groupby": [c_int, c_vp, c_int, c_vp, c_u64, c_vp, ctypes.c_uint32, c_u64,
                            ctypes.c_uint32, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                            c_vp, c_u64, c_vp],
    "jfx_batched_key_ranges": [c_int, c_vp, c_vp, c_vp, ctypes.c_uint32, c_vp, c_vp],
    "jfx_batched_routed_workspace_size": [ctypes.c_uint32, c_int, ctypes.c_uint32,
                                          ctypes.POINTER(BatchedRouting), c_int, c_u64p],
    "jfx_batched_groupby_routed": [c_int, c_vp, c_int, c_vp, c_u64, c_vp, ctypes.c_uint32, c_u64,
                                   ctypes.c_uint32, c_int, c_int, ctypes.POINTER(BatchedRouting),
                                   c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_u64, c_vp],
    "jfx_served_join_workspace_size": [c_u64, ctypes.c_uint32, c_u64p],
    "jfx_served_join_count": [c_int, c_vp, c_vp, ctypes.c_uint32, c_vp, c_u64, c_u64p, c_vp, c_u64,
                              c_vp],
    "jfx_served_join_fill": [c_int, c_vp, c_vp, ctypes.c_uint32, c_int, c_vp, c_u64, c_int, c_vp,
                             c_u64, c_vp, c_vp, c_vp, c_vp, c_u64, c_vp],
    "jfx_served_batch_workspace_size": [ctypes.c_uint32, c_u64, c_u64p],
    "jfx_served_batch": [c_int, c_int, c_vp, ctypes.c_uint32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                         c_u64, c_vp],
    "jfx_served_groups": [c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_uint32, c_int, c_vp,
                          c_vp, c_vp, c_vp, c_vp, c_vp],
    "jfx_compact_workspace_size": [c_u64, c_u64p],
    "jfx_compact_rows": [c_vp, c_u64, c_vp, c_u64p, c_vp, c_u64, c_vp],
    "jfx_filter_rows": [c_int, c_vp, c_int, c_vp, c_u64, c_vp, c_u64p, c_vp, c_u64, c_vp],
    "jfx_pack_keys": [ctypes.POINTER(c_vp), ctypes.POINTER(c_int), c_int, c_u64p, c_u32p,
                      ctypes.POINTER(c_int), c_u64, c_vp, c_vp],
    "jfx_unpack_keys": [c_vp, c_u64, ctypes.POINTER(c_int), c_int, c_u64p, c_u32p,
                        ctypes.POINTER(c_int), ctypes.POINTER(c_vp), c_vp],
    "jfx_groupby_workspace_size": [c_int, c_int, c_u64, c_u64p],
    "jfx_groupby": [c_int, c_vp, c_int, c_vp, c_u64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_u64p,
                    c_int, c_vp, c_u64, c_vp],
    "jfx_groupby_ex": [c_int, c_vp, c_int, c_vp, c_u64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_u64p,
                       c_u64, c_int, c_vp, c_u64, c_vp],
    "jfx_bandwidth_probe": [c_int, c_vp, c_vp, c_u64, ctypes.c_uint32, ctypes.c_uint32, c_int,
                            c_vp],
    "jfx_groupby_noalloc": [c_int, c_vp, c_int, c_vp, c_u64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                            c_u64p, c_u64, c_vp],
}


def lib():
    """Load libjfx.so once; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise JfxError(f"{LIB_PATH} is missing: run `python -m jaxframes_amd.build` "
                           "(there is no CPU fallback)")
        L = ctypes.CDLL(str(LIB_PATH))
        L.jfx_last_error.restype = ctypes.c_char_p
        L.jfx_last_error.argtypes = []
        for name, args in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return ["jfx_last_error", *_SIGNATURES.keys()]


def check(status: int, what: str):
    if status != 0:
        msg = lib().jfx_last_error().decode(errors="replace")
        raise JfxError(f"{what}: {_ERRNAMES.get(status, status)}: {msg}")


_inject_lock = threading.Lock()
_inject: dict = {}   # entry-point name prefix -> [status, remaining calls]


def inject_failure(prefix: str, status: int = 2, count: int = 1):
    """Test hook: the next `count` calls whose entry-point name starts with
    prefix fail with `status` (a JFX_ERR_* code) without running, as if the
    library had returned it. count 0 removes the hook."""
    with _inject_lock:
        if count <= 0:
            _inject.pop(prefix, None)
        else:
            _inject[prefix] = [status, count]


def call(name: str, *args):
    if _inject:
        with _inject_lock:
            for prefix, slot in list(_inject.items()):
                if name.startswith(prefix):
                    slot[1] -= 1
                    if slot[1] <= 0:
                        del _inject[prefix]
                    raise JfxError(f"{name}: {_ERRNAMES.get(slot[0], slot[0])}: injected failure")
    check(getattr(lib(), name)(*args), name)


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream_handle(device=None) -> int:
    """The current HIP stream of device (None: the current device) as an
    integer handle (torch's raw-stream query, a tenth of the cost of
    torch.cuda.current_stream, where this torch has it)."""
    if _raw_stream is None:
        return torch.cuda.current_stream(device).cuda_stream
    if device is None:
        idx = torch.cuda.current_device()
    elif isinstance(device, int):
        idx = device
    elif isinstance(device, torch.device) and device.index is not None:
        idx = device.index
    else:
        idx = torch.device(device).index
        if idx is None:
            idx = torch.cuda.current_device()
    return _raw_stream(idx)


_gpu_seen = False


def require_gpu():
    global _gpu_seen
    if _gpu_seen:
        return
    if not torch.cuda.is_available():
        raise JfxError("no ROCm GPU is visible; jaxframes_amd runs on MI355X only")
    _gpu_seen = True


_ws_cache: dict = {}


def workspace(nbytes: int, device: torch.device, stream: int | None = None) -> torch.Tensor:
    """A cached device byte buffer of at least nbytes for (device, stream);
    stream is the current stream's handle when the caller already holds it."""
    key = (device.index if device.index is not None else torch.cuda.current_device(),
           stream_handle(device) if stream is None else stream)
    buf = _ws_cache.get(key)
    if buf is None or buf.numel() < nbytes:
        cap = max(int(nbytes * 1.25), 1 << 20)
        buf = torch.empty(cap, dtype=torch.uint8, device=device)
        _ws_cache[key] = buf
    return buf


def device_error() -> int:
    code = ctypes.c_uint32(0)
    call("jfx_device_error", c_vp(stream_handle()), ctypes.byref(code))
    return code.value


def ptr(t: torch.Tensor | None) -> c_vp:
    return c_vp(0 if t is None else t.data_ptr())


class KernelTimer:
    """Per-kernel HIP-event timing inside libjfx (events recorded on the
    stream each kernel is launched on). Use as a context manager around the
    timed region, then read .stats() -> {name: (total_ms, launches)}.
    prefix limits the timing to kernels whose name starts with it, every to
    one in `every` of those launches (launches counts the timed ones)."""

    def __init__(self, prefix: str = "", every: int = 1):
        self.prefix = prefix
        self.every = every

    def __enter__(self):
        call("jfx_prof_reset")
        call("jfx_prof_filter", self.prefix.encode())
        call("jfx_prof_sample", self.every)
        call("jfx_prof_enable", 1)
        return self

    def __exit__(self, *exc):
        call("jfx_prof_enable", 0)
        call("jfx_prof_filter", b"")
        call("jfx_prof_sample", 1)

    @staticmethod
    def stats() -> dict:
        n = c_int(0)
        call("jfx_prof_count", ctypes.byref(n))
        out = {}
        buf = ctypes.create_string_buffer(256)
        for i in range(n.value):
            ms = ctypes.c_double(0.0)
            cnt = ctypes.c_uint64(0)
            call("jfx_prof_get", i, buf, 256, ctypes.byref(ms), ctypes.byref(cnt))
            out[buf.value.decode()] = (ms.value, cnt.value)
        return out
