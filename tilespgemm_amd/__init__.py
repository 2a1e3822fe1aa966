"""tilespgemm_amd -- MI355X (gfx950) tiled FP64 SpGEMM behind a C ABI.

Python host side over libtsg.so (ctypes).  torch is used only for device
memory and streams; every computation runs in the library's HIP kernels, and
the functions here fail loudly when libtsg.so or a GPU is missing -- there is
no CPU fallback.

    >>> import tilespgemm_amd as tsg
    >>> h = tsg.Handle()                       # device 0
    >>> ic, jc, c = h.spgemm(A, B)             # A, B: tsg.DeviceCsr on cuda:0

Host-pointer, reference-compatible entry (thmkl_dcsrmultcsr.c:7-10):

    >>> ic = tsg.thmkl_dcsrmultcsr(1, m, n, k, A, B)                # request=1
    >>> jc, c = tsg.thmkl_dcsrmultcsr(0, m, n, k, A, B, ic=ic)      # request=0
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtsg.so")

METHOD_AUTO, METHOD_TILE, METHOD_ROW = 0, 1, 2
FP_FUSED, FP_SEPARATE = 0, 1

ERRORS = {
    -1: "TSG_ERR_INVALID", -2: "TSG_ERR_HIP", -3: "TSG_ERR_ALLOC",
    -4: "TSG_ERR_OVERFLOW", -5: "TSG_ERR_STATE", -6: "TSG_ERR_NODEVICE",
}

# Plan arrays (include/tsg.h), name -> (enum, numpy dtype)
PLAN_ARRAYS = {
    "A_TILE_PTR": (0, np.int32), "A_TILE_COL": (1, np.int32), "A_TILE_OFF": (2, np.int32),
    "A_MASK": (3, np.uint16), "A_ROWPTR": (4, np.uint8), "A_IDX": (5, np.uint8),
    "A_VAL": (6, np.float64),
    "B_TILE_PTR": (7, np.int32), "B_TILE_COL": (8, np.int32), "B_TILE_OFF": (9, np.int32),
    "B_MASK": (10, np.uint16), "B_ROWPTR": (11, np.uint8), "B_IDX": (12, np.uint8),
    "B_VAL": (13, np.float64),
    "B_CSC_PTR": (14, np.int32), "B_CSC_ROW": (15, np.int32), "B_CSC_TILE": (16, np.int32),
    "C_TILE_PTR": (17, np.int32), "C_TILE_COL": (18, np.int32), "C_MASK": (19, np.uint16),
    "C_ROWPTR": (20, np.uint8), "C_TILE_OFF": (21, np.int64),
}

# Every symbol include/tsg.h declares, with its ctypes signature.
_vp, _i, _i64, _sz = C.c_void_p, C.c_int, C.c_int64, C.c_size_t
SIGNATURES = {
    "thmkl_dcsrmultcsr": (None, [_vp] * 17),
    "tsg_dcsrmultcsr_i64": (_i, [_i, _i, _i, _i] + [_vp] * 9),
    "tsg_set_host_slab_entries": (_i, [_i64]),
    "tsg_get_host_slab_entries": (_i64, []),
    "tsg_host_last_slabs": (_i, []),
    "tsg_create": (_i, [_vp, _i]),
    "tsg_destroy": (_i, [_vp]),
    "tsg_set_stream": (_i, [_vp, _vp]),
    "tsg_set_method": (_i, [_vp, _i]),
    "tsg_set_fp_mode": (_i, [_vp, _i]),
    "tsg_get_method_used": (_i, [_vp]),
    "tsg_workspace_bytes": (_sz, [_vp]),
    "tsg_spgemm_symbolic": (_i, [_vp, _i, _i, _i, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _vp]),
    "tsg_spgemm_numeric": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _i64,
                                _vp, _vp, _vp]),
    "tsg_spgemm_noalloc": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _i64,
                                _vp, _vp, _vp, _i64, _vp]),
    "tsg_batch_symbolic": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "tsg_batch_numeric": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "tsg_set_tile_density": (_i, [_vp, C.c_double]),
    "tsg_set_row_bins": (_i, [_vp, _i, _i, _i]),
    "tsg_get_tuning": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "tsg_set_item_order": (_i, [_vp, _i64]),
    "tsg_get_item_order": (_i, [_vp, _vp]),
    "tsg_set_bitmap_budget": (_i, [_vp, _i64]),
    "tsg_set_block_pointer_budget": (_i, [_vp, _i64]),
    "tsg_read_mtx": (_i, [C.c_char_p, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "tsg_free": (None, [_vp]),
    "tsg_server_create": (_i, [_vp, _i, _vp]),
    "tsg_server_submit": (_i, [_vp, _vp, _vp]),
    "tsg_server_wait": (_i, [_vp, C.c_uint64, _vp]),
    "tsg_server_release": (_i, [_vp, C.c_uint64]),
    "tsg_server_destroy": (_i, [_vp]),
    "tsg_csr_transpose": (_i, [_vp, _i, _i, _vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "tsg_intermediate_products": (_i, [_vp, _i, _i, _vp, _vp, _i64, _vp, _vp]),
    "tsg_plan_array_bytes": (_sz, [_vp, _i]),
    "tsg_plan_array_copy": (_i, [_vp, _i, _vp, _sz]),
    "tsg_enable_timing": (_i, [_vp, _i]),
    "tsg_stage_count": (_i, [_vp]),
    "tsg_stage_name": (C.c_char_p, [_vp, _i]),
    "tsg_stage_ms": (C.c_double, [_vp, _i]),
    "tsg_last_error": (C.c_char_p, [_vp]),
}

_lib = None


class TsgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def lib_path() -> str:
    return os.environ.get("TSG_LIB", LIB_PATH)


def load():
    """Load libtsg.so (built by tilespgemm_amd/build.py); raise if absent."""
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"{path} not found: build it with `python -m tilespgemm_amd.build`")
        # torch ships its own HIP runtime under ROCm's soname
        # (libamdhip64.so.7): with torch loaded first, libtsg binds to that
        # runtime and one runtime serves both; loaded the other way round, the
        # process would hold two runtimes, and torch's could not use the GPU
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def compile_library() -> str:
    """Build libtsg.so in-tree (tilespgemm_amd/build.py)."""
    import importlib
    return importlib.import_module("tilespgemm_amd.build").build()


def _check(rc: int, handle=None):
    if rc != 0:
        msg = load().tsg_last_error(handle)
        raise TsgError(rc, msg.decode() if msg else "")


# ------------------------------------------------------------ device CSR
@dataclass
class DeviceCsr:
    """CSR matrix resident on the GPU (torch tensors: int32 indptr/indices,
    float64 data)."""
    rows: int
    cols: int
    indptr: "object"
    indices: "object"
    data: "object"

    @property
    def nnz(self) -> int:
        return int(self.indices.numel())

    @classmethod
    def from_host(cls, rows, cols, indptr, indices, data, device="cuda:0"):
        import torch
        indptr = np.asarray(indptr)
        if indptr[-1] >= 2**31:
            raise TsgError(-4, "input nnz must fit int32")
        return cls(rows, cols,
                   torch.from_numpy(np.ascontiguousarray(indptr, dtype=np.int32)).to(device),
                   torch.from_numpy(np.ascontiguousarray(indices, dtype=np.int32)).to(device),
                   torch.from_numpy(np.ascontiguousarray(data, dtype=np.float64)).to(device))


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


class Product(C.Structure):
    """tsg_product_t (include/tsg.h): one product C_i = A_i * B_i of a batch."""
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("k", C.c_int),
                ("ia", C.c_void_p), ("ja", C.c_void_p), ("a", C.c_void_p), ("nnz_a", C.c_int64),
                ("ib", C.c_void_p), ("jb", C.c_void_p), ("b", C.c_void_p), ("nnz_b", C.c_int64)]

    @classmethod
    def of(cls, A: "DeviceCsr", B: "DeviceCsr"):
        return cls(A.rows, A.cols, B.cols, A.indptr.data_ptr(), A.indices.data_ptr(),
                   A.data.data_ptr(), A.nnz, B.indptr.data_ptr(), B.indices.data_ptr(),
                   B.data.data_ptr(), B.nnz)


MTX_ERRORS = {-101: "cannot open", -102: "bad banner", -104: "bad size line",
              -105: "short file or index out of range"}


def read_mtx(path: str):
    """Matrix Market file -> (m, n, indptr int32, indices int32, data float64,
    symmetric) with the reference loader's semantics (tsg_read_mtx)."""
    m, n, sym = C.c_int(0), C.c_int(0), C.c_int(0)
    nnz = C.c_int64(0)
    pia, pja, pa = C.c_void_p(), C.c_void_p(), C.c_void_p()
    lib = load()
    rc = lib.tsg_read_mtx(os.fsencode(path), C.byref(m), C.byref(n), C.byref(nnz), C.byref(sym),
                          C.byref(pia), C.byref(pja), C.byref(pa))
    if rc != 0:
        raise TsgError(rc, f"{path}: {MTX_ERRORS.get(rc, 'read failed')}")
    try:
        k = nnz.value
        ip = np.ctypeslib.as_array(C.cast(pia, C.POINTER(C.c_int)), (m.value + 1,)).copy()
        ix = np.ctypeslib.as_array(C.cast(pja, C.POINTER(C.c_int)), (max(k, 1),))[:k].copy()
        dv = np.ctypeslib.as_array(C.cast(pa, C.POINTER(C.c_double)), (max(k, 1),))[:k].copy()
    finally:
        for ptr in (pia, pja, pa):
            lib.tsg_free(ptr)
    return m.value, n.value, ip, ix, dv, bool(sym.value)


class ServerConfig(C.Structure):
    """tsg_server_config_t (include/tsg.h)."""
    _fields_ = [("workers", C.c_int), ("max_batch", C.c_int), ("mem_budget", C.c_size_t),
                ("method", C.c_int), ("tile_density", C.c_double), ("bin1", C.c_int),
                ("bin2", C.c_int), ("bin3", C.c_int)]


class Result(C.Structure):
    """tsg_result_t (include/tsg.h)."""
    _fields_ = [("status", C.c_int), ("m", C.c_int), ("nnz", C.c_int64),
                ("ic", C.c_void_p), ("jc", C.c_void_p), ("c", C.c_void_p),
                ("queue_ms", C.c_double), ("total_ms", C.c_double), ("batch_size", C.c_int),
                ("products", C.c_int64)]


class Server:
    """Continuous-batching request server on one device (tsg_server_*).

    submit(A, B) enqueues C = A*B on device-resident operands (keep them alive
    until the result is released) and returns a ticket; wait(ticket) blocks
    and returns the Result; fetch(ticket) copies it to host numpy arrays and
    releases it.  The estimator and each worker use a stream of their own:
    set GPU_MAX_HW_QUEUES to at least workers + 1 plus the caller's busy
    streams before HIP starts (its default is 4), or streams that share a
    hardware queue run in turn."""

    def __init__(self, device: int = 0, workers: int = 2, max_batch: int = 64,
                 mem_budget: int = 32 << 30, method: int = METHOD_AUTO, tuning=None):
        self._s = C.c_void_p()
        t = tuning or {}
        bins = t.get("row_bins", (0, 0, 0))
        cfg = ServerConfig(workers, max_batch, mem_budget, method,
                           float(t.get("tile_density", 0.0)), *[int(x) for x in bins])
        _check(load().tsg_server_create(C.byref(self._s), device, C.byref(cfg)))
        self._keep = {}

    def submit(self, A: DeviceCsr, B: DeviceCsr) -> int:
        req = Product.of(A, B)
        t = C.c_uint64(0)
        _check(load().tsg_server_submit(self._s, C.byref(req), C.byref(t)))
        self._keep[t.value] = (A, B)
        return t.value

    def wait(self, ticket: int) -> Result:
        res = Result()
        rc = load().tsg_server_wait(self._s, ticket, C.byref(res))
        if rc != 0:
            raise TsgError(rc, f"request {ticket} failed")
        return res

    def release(self, ticket: int):
        _check(load().tsg_server_release(self._s, ticket))
        self._keep.pop(ticket, None)

    def fetch(self, ticket: int):
        """(ic int64, jc int32, c float64) on the host, then release."""
        import torch
        res = self.wait(ticket)
        ic = np.empty(res.m + 1, dtype=np.int64)
        jc = np.empty(res.nnz, dtype=np.int32)
        c = np.empty(res.nnz, dtype=np.float64)
        torch.cuda.synchronize()
        _d2h(ic, res.ic)
        if res.nnz:
            _d2h(jc, res.jc)
            _d2h(c, res.c)
        self.release(ticket)
        return ic, jc, c, res

    def fetch_device(self, ticket: int):
        """(ic int64, jc int32, c float64) as new torch tensors on the
        server's device (device-to-device copies), then release."""
        import torch
        res = self.wait(ticket)
        dev = torch.device("cuda", torch.cuda.current_device())
        ic = torch.empty(res.m + 1, dtype=torch.int64, device=dev)
        jc = torch.empty(max(res.nnz, 1), dtype=torch.int32, device=dev)
        c = torch.empty(max(res.nnz, 1), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        _d2d(ic, res.ic)
        if res.nnz:
            _d2d(jc, res.jc)
            _d2d(c, res.c)
        self.release(ticket)
        return ic, jc[:res.nnz], c[:res.nnz], res

    def close(self):
        if getattr(self, "_s", None):
            load().tsg_server_destroy(self._s)
            self._s = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _d2h(dst: np.ndarray, src_ptr: int):
    """Copy device memory at src_ptr into the host array dst with the HIP
    runtime libtsg itself is linked against (symbol lookup through the
    library's handle searches its dependencies)."""
    f = load().hipMemcpy
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rc = f(dst.ctypes.data_as(C.c_void_p), C.c_void_p(src_ptr), dst.nbytes, 2)
    if rc != 0:
        raise TsgError(-5, f"hipMemcpy device->host failed ({rc})")


def _d2d(dst, src_ptr: int):
    """Copy device memory at src_ptr into the torch tensor dst (same
    device; synchronous, through libtsg's HIP runtime)."""
    f = load().hipMemcpy
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rc = f(C.c_void_p(dst.data_ptr()), C.c_void_p(src_ptr),
           dst.numel() * dst.element_size(), 3)
    if rc != 0:
        raise TsgError(-5, f"hipMemcpy device->device failed ({rc})")


class Handle:
    """A tsg handle on one device.  Work is issued on torch's current stream
    of that device (or, for torch's default stream, on a blocking stream that
    orders with it), so results order naturally with torch operations."""

    def __init__(self, device: int = 0, method: int = METHOD_AUTO,
                 fp_mode: int = FP_FUSED, timing: bool = False):
        import torch
        self._torch = torch
        self.device = device
        h = C.c_void_p()
        _check(load().tsg_create(C.byref(h), device))
        self._h = h
        with torch.cuda.device(device):
            self.stream = torch.cuda.current_stream()
        # torch's default stream is the null stream; the handle's own stream is
        # a blocking stream that orders with it.  Any other torch stream is
        # used directly.
        if self.stream.cuda_stream:
            _check(load().tsg_set_stream(self._h, C.c_void_p(self.stream.cuda_stream)), self._h)
        self.set_method(method)
        _check(load().tsg_set_fp_mode(self._h, fp_mode), self._h)
        if timing:
            _check(load().tsg_enable_timing(self._h, 1), self._h)

    def close(self):
        if getattr(self, "_h", None):
            load().tsg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        """Issue work on a torch stream (None: the handle's own stream)."""
        ptr = stream.cuda_stream if stream is not None else 0
        _check(load().tsg_set_stream(self._h, C.c_void_p(ptr)), self._h)
        self.stream = stream

    def set_method(self, method: int):
        _check(load().tsg_set_method(self._h, method), self._h)

    def set_tuning(self, tuning: dict):
        """Apply {"tile_density": x, "row_bins": (b1, b2, b3),
        "block_pointer_budget": bytes} (every key optional; see
        tsg_set_tile_density / tsg_set_row_bins / tsg_set_block_pointer_budget)."""
        if "tile_density" in tuning:
            _check(load().tsg_set_tile_density(self._h, float(tuning["tile_density"])), self._h)
        if "row_bins" in tuning:
            b1, b2, b3 = (int(x) for x in tuning["row_bins"])
            _check(load().tsg_set_row_bins(self._h, b1, b2, b3), self._h)
        if "block_pointer_budget" in tuning:
            self.set_block_pointer_budget(int(tuning["block_pointer_budget"]))

    def tuning(self) -> dict:
        d = C.c_double(0)
        b = [C.c_int(0) for _ in range(3)]
        _check(load().tsg_get_tuning(self._h, C.byref(d), *[C.byref(x) for x in b]), self._h)
        return {"tile_density": d.value, "row_bins": tuple(x.value for x in b)}

    def set_item_order(self, min_items: int):
        """Big-row items run in first-column order from min_items items on
        (tsg_set_item_order)."""
        _check(load().tsg_set_item_order(self._h, int(min_items)), self._h)

    def set_bitmap_budget(self, nbytes: int):
        """Bytes reserved for the big rows' column bitmaps (sparse form)
        kept for the rank-indexed numeric kernel (tsg_set_bitmap_budget;
        negative: automatic)."""
        _check(load().tsg_set_bitmap_budget(self._h, int(nbytes)), self._h)

    def set_block_pointer_budget(self, nbytes: int):
        """Bytes of the big-row B block-pointer table before blocks widen
        (tsg_set_block_pointer_budget; negative: automatic)."""
        _check(load().tsg_set_block_pointer_budget(self._h, int(nbytes)), self._h)

    def item_order(self) -> int:
        v = C.c_int64(0)
        _check(load().tsg_get_item_order(self._h, C.byref(v)), self._h)
        return v.value

    def set_fp_mode(self, fp_mode: int):
        _check(load().tsg_set_fp_mode(self._h, fp_mode), self._h)

    @property
    def method_used(self) -> int:
        return load().tsg_get_method_used(self._h)

    @property
    def workspace_bytes(self) -> int:
        return int(load().tsg_workspace_bytes(self._h))

    def symbolic(self, A: DeviceCsr, B: DeviceCsr, ic=None):
        """Row pointers of C = A*B (int64, m+1) and nnz(C)."""
        torch = self._torch
        if A.cols != B.rows:
            raise TsgError(-1, "A.cols != B.rows")
        if ic is None:
            ic = torch.empty(A.rows + 1, dtype=torch.int64, device=A.indptr.device)
        nnz = C.c_int64(0)
        _check(load().tsg_spgemm_symbolic(
            self._h, A.rows, A.cols, B.cols, _ptr(A.indptr), _ptr(A.indices), A.nnz,
            _ptr(B.indptr), _ptr(B.indices), B.nnz, _ptr(ic), C.byref(nnz)), self._h)
        return ic, int(nnz.value)

    def numeric(self, A: DeviceCsr, B: DeviceCsr, ic, jc, c):
        _check(load().tsg_spgemm_numeric(
            self._h, A.rows, A.cols, B.cols, _ptr(A.indptr), _ptr(A.indices), _ptr(A.data),
            A.nnz, _ptr(B.indptr), _ptr(B.indices), _ptr(B.data), B.nnz, _ptr(ic), _ptr(jc),
            _ptr(c)), self._h)

    def spgemm(self, A: DeviceCsr, B: DeviceCsr):
        """C = A*B on the device; returns (ic int64, jc int32, c float64)."""
        torch = self._torch
        ic, nnz = self.symbolic(A, B)
        dev = A.indptr.device
        jc = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
        c = torch.empty(max(nnz, 1), dtype=torch.float64, device=dev)
        self.numeric(A, B, ic, jc, c)
        return ic, jc[:nnz], c[:nnz]

    def spgemm_noalloc(self, A: DeviceCsr, B: DeviceCsr, ic, jc, c) -> int:
        """C = A*B into caller-allocated tensors (ic: m+1 int64; jc, c: room
        for nzmax = jc.numel() entries) with no allocation inside the call
        (tsg_spgemm_noalloc).  Returns nnz(C); raises TsgError when C does not
        fit or the product's workspace exceeds the library's scratch."""
        if A.cols != B.rows:
            raise TsgError(-1, "A.cols != B.rows")
        nzmax = min(int(jc.numel()), int(c.numel()))
        nnz = C.c_int64(0)
        _check(load().tsg_spgemm_noalloc(
            self._h, A.rows, A.cols, B.cols, _ptr(A.indptr), _ptr(A.indices), _ptr(A.data),
            A.nnz, _ptr(B.indptr), _ptr(B.indices), _ptr(B.data), B.nnz, _ptr(ic), _ptr(jc),
            _ptr(c), nzmax, C.byref(nnz)), self._h)
        return int(nnz.value)

    # ---------------------------------------------------------- batches
    def batch_symbolic(self, pairs):
        """pairs: list of (A_i, B_i) DeviceCsr.  Returns (products array, ic
        (sum m_i + 1, int64, device), nnz per product (host int64 array),
        row offsets per product)."""
        torch = self._torch
        arr = (Product * len(pairs))(*[Product.of(A, B) for A, B in pairs])
        rows = sum(A.rows for A, _ in pairs)
        dev = pairs[0][0].indptr.device
        ic = torch.empty(rows + 1, dtype=torch.int64, device=dev)
        tot = C.c_int64(0)
        per = np.zeros(len(pairs), dtype=np.int64)
        _check(load().tsg_batch_symbolic(self._h, len(pairs), arr, _ptr(ic), C.byref(tot),
                                         per.ctypes.data_as(C.c_void_p)), self._h)
        roff = np.concatenate([[0], np.cumsum([A.rows for A, _ in pairs])])
        return arr, ic, per, roff, int(tot.value)

    def batch_numeric(self, arr, ic, jc, c):
        _check(load().tsg_batch_numeric(self._h, len(arr), arr, _ptr(ic), _ptr(jc), _ptr(c)),
               self._h)

    def batch_spgemm(self, pairs):
        """Batch of C_i = A_i * B_i; returns a list of (ic_i, jc_i, c_i) device
        tensors (ic_i int64 starting at 0, local columns)."""
        torch = self._torch
        arr, ic, per, roff, tot = self.batch_symbolic(pairs)
        dev = ic.device
        jc = torch.empty(max(tot, 1), dtype=torch.int32, device=dev)
        c = torch.empty(max(tot, 1), dtype=torch.float64, device=dev)
        self.batch_numeric(arr, ic, jc, c)
        out = []
        for i in range(len(pairs)):
            r0, r1 = int(roff[i]), int(roff[i + 1])
            sub = ic[r0:r1 + 1]
            e0 = int(sub[0]) if r1 >= r0 else 0
            e1 = e0 + int(per[i])
            out.append((sub - e0, jc[e0:e1], c[e0:e1]))
        return out

    def transpose(self, A: DeviceCsr) -> DeviceCsr:
        """A^T as a DeviceCsr (rows by ascending source row)."""
        torch = self._torch
        dev = A.indptr.device
        tp = torch.empty(A.cols + 1, dtype=torch.int32, device=dev)
        ti = torch.empty(max(A.nnz, 1), dtype=torch.int32, device=dev)
        tv = torch.empty(max(A.nnz, 1), dtype=torch.float64, device=dev)
        _check(load().tsg_csr_transpose(self._h, A.rows, A.cols, _ptr(A.indptr),
                                        _ptr(A.indices), _ptr(A.data), A.nnz, _ptr(tp),
                                        _ptr(ti), _ptr(tv)), self._h)
        return DeviceCsr(A.cols, A.rows, tp, ti[:A.nnz], tv[:A.nnz])

    def intermediate_products(self, A: DeviceCsr, B: DeviceCsr) -> int:
        out = C.c_uint64(0)
        if A.cols != B.rows:
            raise TsgError(-1, "A.cols != B.rows")
        _check(load().tsg_intermediate_products(self._h, A.rows, A.cols, _ptr(A.indptr),
                                                _ptr(A.indices), A.nnz, _ptr(B.indptr),
                                                C.byref(out)), self._h)
        return int(out.value)

    def plan_array(self, name: str) -> np.ndarray:
        which, dt = PLAN_ARRAYS[name]
        nbytes = int(load().tsg_plan_array_bytes(self._h, which))
        out = np.empty(nbytes // np.dtype(dt).itemsize, dtype=dt)
        if nbytes:
            _check(load().tsg_plan_array_copy(self._h, which, out.ctypes.data_as(C.c_void_p),
                                              nbytes), self._h)
        return out

    def stage_times(self) -> dict:
        L = load()
        n = L.tsg_stage_count(self._h)
        out = {}
        for i in range(n):
            name = L.tsg_stage_name(self._h, i).decode()
            out[name] = out.get(name, 0.0) + L.tsg_stage_ms(self._h, i)
        return out


# ----------------------------------------------------- host-pointer entries
def thmkl_dcsrmultcsr(request: int, m: int, n: int, k: int, A, B, ic=None):
    """Reference-compatible two-phase call on host arrays (int32 ic).

    request=1 -> returns ic (int32, m+1); request=0 -> returns (jc, c).
    Raises TsgError when *info < 0."""
    L = load()
    ia = np.ascontiguousarray(A[0], dtype=np.int32)
    ja = np.ascontiguousarray(A[1], dtype=np.int32)
    a = np.ascontiguousarray(A[2], dtype=np.float64)
    same = B is A
    ib = ia if same else np.ascontiguousarray(B[0], dtype=np.int32)
    jb = ja if same else np.ascontiguousarray(B[1], dtype=np.int32)
    b = a if same else np.ascontiguousarray(B[2], dtype=np.float64)
    cm, cn, ck = C.c_int(m), C.c_int(n), C.c_int(k)
    req, srt, nzmax, info = C.c_int(request), C.c_int(1), C.c_int(0), C.c_int(0)
    P = lambda x: x.ctypes.data_as(C.c_void_p)  # noqa: E731
    if request == 1:
        ic = np.zeros(m + 1, dtype=np.int32)
        jc = np.zeros(1, dtype=np.int32)
        c = np.zeros(1, dtype=np.float64)
    else:
        ic = np.ascontiguousarray(ic, dtype=np.int32)
        nnz = int(ic[-1])
        jc = np.zeros(max(nnz, 1), dtype=np.int32)
        c = np.zeros(max(nnz, 1), dtype=np.float64)
    L.thmkl_dcsrmultcsr(b"n", C.byref(req), C.byref(srt), C.byref(cm), C.byref(cn),
                        C.byref(ck), P(a), P(ja), P(ia), P(b), P(jb), P(ib), P(c), P(jc),
                        P(ic), C.byref(nzmax), C.byref(info))
    if info.value < 0:
        raise TsgError(info.value, (L.tsg_last_error(None) or b"").decode())
    if request == 1:
        return ic
    nnz = int(ic[-1])
    return jc[:nnz], c[:nnz]


def dcsrmultcsr_i64(request: int, m: int, n: int, k: int, A, B, ic=None):
    """As thmkl_dcsrmultcsr with an int64 ic (nnz(C) may exceed 2^31-1)."""
    L = load()
    ia = np.ascontiguousarray(A[0], dtype=np.int32)
    ja = np.ascontiguousarray(A[1], dtype=np.int32)
    a = np.ascontiguousarray(A[2], dtype=np.float64)
    same = B is A
    ib = ia if same else np.ascontiguousarray(B[0], dtype=np.int32)
    jb = ja if same else np.ascontiguousarray(B[1], dtype=np.int32)
    b = a if same else np.ascontiguousarray(B[2], dtype=np.float64)
    P = lambda x: x.ctypes.data_as(C.c_void_p)  # noqa: E731
    if request == 1:
        ic = np.zeros(m + 1, dtype=np.int64)
        jc = np.zeros(1, dtype=np.int32)
        c = np.zeros(1, dtype=np.float64)
    else:
        ic = np.ascontiguousarray(ic, dtype=np.int64)
        nnz = int(ic[-1])
        jc = np.zeros(max(nnz, 1), dtype=np.int32)
        c = np.zeros(max(nnz, 1), dtype=np.float64)
    rc = L.tsg_dcsrmultcsr_i64(request, m, n, k, P(a), P(ja), P(ia), P(b), P(jb), P(ib),
                               P(c), P(jc), P(ic))
    _check(rc)
    if request == 1:
        return ic
    nnz = int(ic[-1])
    return jc[:nnz], c[:nnz]


def set_host_slab_entries(entries: int) -> None:
    """Cap on the C entries the host entries' request=0 call keeps on the
    device at once; larger products stream C in row slabs (0: automatic)."""
    _check(load().tsg_set_host_slab_entries(int(entries)))


def host_slab_entries() -> int:
    return int(load().tsg_get_host_slab_entries())


def host_last_slabs() -> int:
    """Row slabs the last request=0 host call ran (1: C whole on the device)."""
    return int(load().tsg_host_last_slabs())
