"""fishfry_amd -- MI355X (gfx950) periodic Poisson solver behind fishfry's
ParisPencil interface.

The compute path is libfishfry.so (hand-written HIP kernels, built in-tree by
``python -c "import __graft_entry__; __graft_entry__.build()"`` or
``make -C fishfry_amd/csrc``).  This module binds its C-ABI (include/fishfry.h)
with ctypes and mirrors the reference's plugin interface:

* ``ParisPencil(n, lo, hi, m, id)`` with ``bytes()`` and
  ``solve(bytes, density, potential, stamps)`` -- ParisPencil.hpp:23-44;
* ``Field`` / ``field_init`` / ``field_diff`` -- FishFry.hpp:14-101;
* ``FishFry(n_tasks, n_points).run(n_iters)`` -- FishFry.hpp:103-229.

Device buffers are torch CUDA tensors (PyTorch is used only for device memory
and streams) or raw device pointers.  There is no CPU fallback: if the library
is missing or fails, calls raise.
"""
from __future__ import annotations

import ctypes
import math
import os
from typing import List, Optional, Sequence, Tuple

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfishfry.so")

FF_OK = 0
FF_ERR_INVALID = 1
FF_ERR_UNSUPPORTED = 2
FF_ERR_BYTES = 3
FF_ERR_HIP = 4
FF_ERR_NOMEM = 5
FF_ERR_REJECTED = 6
FF_ERR_TIMEOUT = 7
FF_ERR_NOT_FOUND = 8

FF_FP64 = 0
FF_FP32_REFINE = 1
FF_PRESERVE_INPUT = 1
FF_BLOCKED_LAYOUT = 2
FF_SLAB_PIPELINE = 4
FF_SYMBOL_3PT = 8
FF_SYMBOL_5PT = 16
FF_PENCIL_LAYOUT = 32
FF_PLAIN_LAYOUT = 64
FF_TUNE_SPLIT_LDS = 1
FF_TUNE_TWO_GROUPS = 2
FF_TUNE_GENERIC = 4
FF_TUNE_LDS_TWIDDLES = 8
FF_SOLVE_STAMPS = 14
FF_MAX_MODES = 32

STATUS_NAMES = {
    0: "FF_OK", 1: "FF_ERR_INVALID", 2: "FF_ERR_UNSUPPORTED", 3: "FF_ERR_BYTES", 4: "FF_ERR_HIP",
    5: "FF_ERR_NOMEM", 6: "FF_ERR_REJECTED", 7: "FF_ERR_TIMEOUT", 8: "FF_ERR_NOT_FOUND",
}


class FishFryError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("refine_steps", ctypes.c_int), ("flags", ctypes.c_int),
                ("slab_planes", ctypes.c_int)]


class Stamp(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("seconds", ctypes.c_double)]


class Tune(ctypes.Structure):
    """ff_tune: per-pass override of elements per thread (E) and lines per workgroup (T)."""
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_int), ("precision", ctypes.c_int), ("E", ctypes.c_int),
                ("T", ctypes.c_int), ("flags", ctypes.c_int)]


class BatchItem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3),
                ("precision", ctypes.c_int), ("refine_steps", ctypes.c_int)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("streams", ctypes.c_int), ("max_batch", ctypes.c_int),
                ("batch_window_us", ctypes.c_double), ("flags", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("precision", ctypes.c_int), ("refine_steps", ctypes.c_int),
                ("field", ctypes.c_void_p), ("f_host", ctypes.c_void_p), ("phi_host", ctypes.c_void_p),
                ("want_norms", ctypes.c_int)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("latency_s", ctypes.c_double), ("queue_s", ctypes.c_double),
                ("norms", ctypes.c_double * 3), ("points", ctypes.c_longlong)]


class ServingStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_longlong), ("rejected", ctypes.c_longlong), ("groups", ctypes.c_longlong),
                ("p50_s", ctypes.c_double), ("p99_s", ctypes.c_double), ("points_per_s", ctypes.c_double),
                ("arena_bytes", ctypes.c_size_t), ("arena_peak_bytes", ctypes.c_size_t),
                ("grouped_solves", ctypes.c_longlong)]


FF_SERVE_LARGE_ALONE = 1

FF_BATCH_GRAPH = 1
FF_BATCH_PERSISTENT = 4
FF_BATCH_GROUPED = 8
FF_BATCH_STEPS = 16


class Field(ctypes.Structure):
    """ff_field: grid, phase shift, amplitude and extra modes of f."""
    _fields_ = [
        ("n", ctypes.c_int * 3),
        ("lo", ctypes.c_double * 3),
        ("d", ctypes.c_double * 3),
        ("shift", ctypes.c_double * 3),
        ("scale", ctypes.c_double),
        ("dd", ctypes.c_double * 3),
        ("nmodes", ctypes.c_int),
        ("mode_k", (ctypes.c_int * 3) * FF_MAX_MODES),
        ("mode_a", ctypes.c_double * FF_MAX_MODES),
        ("mode_b", ctypes.c_double * FF_MAX_MODES),
    ]


_lib = None


def lib():
    """Load libfishfry.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FishFryError(FF_ERR_INVALID, f"{LIB_PATH} is missing: build it first (make -C fishfry_amd/csrc)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i3 = ctypes.POINTER(ctypes.c_int)
    d3 = ctypes.POINTER(ctypes.c_double)
    L.ff_last_error.restype = ctypes.c_char_p
    L.ff_last_error.argtypes = []
    L.ff_reference_bytes.restype = ctypes.c_size_t
    L.ff_reference_bytes.argtypes = [i3, i3]
    L.ff_plan_create.restype = ctypes.c_int
    L.ff_plan_create.argtypes = [i3, d3, d3, i3, i3, ctypes.POINTER(Options), ctypes.POINTER(P)]
    L.ff_plan_bytes.restype = ctypes.c_size_t
    L.ff_plan_bytes.argtypes = [P]
    L.ff_plan_destroy.restype = ctypes.c_int
    L.ff_plan_destroy.argtypes = [P]
    L.ff_solve.restype = ctypes.c_int
    L.ff_solve.argtypes = [P, ctypes.c_size_t, P, P, P, ctypes.POINTER(Stamp), ctypes.POINTER(ctypes.c_int)]
    L.ff_solve_async.restype = ctypes.c_int
    L.ff_solve_async.argtypes = [P, ctypes.c_size_t, P, P, P]
    L.ff_stencil_apply.restype = ctypes.c_int
    L.ff_stencil_apply.argtypes = [P, P, P, P]
    L.ff_solve_noalloc.restype = ctypes.c_int
    L.ff_solve_noalloc.argtypes = [P, ctypes.c_size_t, P, P, P]
    L.ff_plan_describe.restype = ctypes.c_int
    L.ff_plan_describe.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t]
    L.ff_solve_pass.restype = ctypes.c_int
    L.ff_solve_pass.argtypes = [P, ctypes.c_int, P, P, P]
    L.ff_tune_set.restype = ctypes.c_int
    L.ff_tune_set.argtypes = [ctypes.POINTER(Tune)]
    L.ff_tune_reset.restype = None
    L.ff_tune_reset.argtypes = []
    L.ff_tune_query.restype = ctypes.c_int
    L.ff_tune_query.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Tune)]
    L.ff_batch_create.restype = ctypes.c_int
    L.ff_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(BatchItem), ctypes.c_int, ctypes.POINTER(P)]
    L.ff_batch_bytes.restype = ctypes.c_size_t
    L.ff_batch_bytes.argtypes = [P, ctypes.c_int]
    L.ff_batch_solve.restype = ctypes.c_int
    L.ff_batch_solve.argtypes = [P, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p), P]
    L.ff_batch_wait.restype = ctypes.c_int
    L.ff_batch_wait.argtypes = [P, P]
    L.ff_batch_autotune.restype = ctypes.c_int
    L.ff_batch_autotune.argtypes = [P, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p), P,
                                    ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]
    L.ff_batch_describe.restype = ctypes.c_int
    L.ff_batch_describe.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t]
    L.ff_batch_destroy.restype = ctypes.c_int
    L.ff_batch_destroy.argtypes = [P]
    L.ff_server_create.restype = ctypes.c_int
    L.ff_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(P)]
    L.ff_server_footprint.restype = ctypes.c_size_t
    L.ff_server_footprint.argtypes = [ctypes.POINTER(Request)]
    L.ff_server_submit.restype = ctypes.c_int
    L.ff_server_submit.argtypes = [P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
    L.ff_server_wait.restype = ctypes.c_int
    L.ff_server_wait.argtypes = [P, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(Result)]
    L.ff_server_stats.restype = ctypes.c_int
    L.ff_server_stats.argtypes = [P, ctypes.POINTER(ServingStats)]
    L.ff_server_destroy.restype = ctypes.c_int
    L.ff_server_destroy.argtypes = [P]
    L.ff_field_default.restype = None
    L.ff_field_default.argtypes = [i3, ctypes.POINTER(Field)]
    L.ff_field_init.restype = ctypes.c_int
    L.ff_field_init.argtypes = [ctypes.POINTER(Field), P, P]
    L.ff_field_diff.restype = ctypes.c_int
    L.ff_field_diff.argtypes = [ctypes.POINTER(Field), P, d3, P]
    _lib = L
    return L


def last_error() -> str:
    return lib().ff_last_error().decode()


def check(status: int, what: str = "") -> None:
    if status != FF_OK:
        raise FishFryError(status, f"{what}: {last_error()}" if what else last_error())


def _i3(v):
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def _d3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


def _ptr(buf) -> int:
    """Device address of a torch tensor or an int."""
    if buf is None:
        return 0
    if isinstance(buf, int):
        return buf
    return int(buf.data_ptr())


def _stream(stream) -> int:
    if stream is None:
        try:
            import torch

            return int(torch.cuda.current_stream().cuda_stream)
        except Exception:
            return 0
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


def reference_bytes(n: Sequence[int], m: Sequence[int] = (1, 1, 1)) -> int:
    """The reference's per-array byte formula (ParisPencil.cpp:49-75)."""
    return int(lib().ff_reference_bytes(_i3(n), _i3(m)))


def benchmark_grid(n: Sequence[int]) -> Tuple[List[float], List[float], List[float]]:
    """lo, hi, dx of FishFry.hpp:109-111,133-134 at 1x1x1 tasks."""
    dx = [2.0 * math.pi / n[a] for a in range(3)]
    return [-math.pi] * 3, [math.pi - dx[a] for a in range(3)], dx


class ParisPencil:
    """Mirror of the reference's solver plugin (ParisPencil.hpp:11-59).

    ``ParisPencil(n, lo, hi, m, id)`` plans the solve on the current device;
    ``bytes()`` is the per-array allocation the caller must make;
    ``solve(bytes, density, potential, stamps)`` overwrites ``potential`` with
    phi (``density`` is clobbered in FP64 mode, as in the reference).
    """

    def __init__(self, n: Sequence[int], lo: Sequence[float], hi: Sequence[float],
                 m: Sequence[int] = (1, 1, 1), id: Sequence[int] = (0, 0, 0),
                 precision: int = FF_FP64, refine_steps: int = 0, flags: int = 0, slab_planes: int = 0):
        self.n = tuple(int(v) for v in n)
        self.precision = precision
        self.refine_steps = refine_steps
        h = ctypes.c_void_p()
        opts = Options(precision, refine_steps, flags, slab_planes)
        check(lib().ff_plan_create(_i3(n), _d3(lo), _d3(hi), _i3(m), _i3(id), ctypes.byref(opts), ctypes.byref(h)),
              "ParisPencil")
        self._h = h

    @property
    def handle(self) -> int:
        return self._h.value or 0

    def bytes(self) -> int:
        return int(lib().ff_plan_bytes(self._h))

    def solve(self, nbytes: int, density, potential, stamps: Optional[list] = None, stream=None) -> None:
        L = lib()
        if stamps is None:
            check(L.ff_solve(self._h, nbytes, _ptr(density), _ptr(potential), _stream(stream), None, None), "solve")
            return
        arr = (Stamp * FF_SOLVE_STAMPS)()
        cnt = ctypes.c_int(FF_SOLVE_STAMPS)
        check(L.ff_solve(self._h, nbytes, _ptr(density), _ptr(potential), _stream(stream), arr, ctypes.byref(cnt)),
              "solve")
        stamps.clear()
        for i in range(cnt.value):
            stamps.append((arr[i].name.decode(), arr[i].seconds))

    def solve_async(self, nbytes: int, density, potential, stream=None) -> None:
        check(lib().ff_solve_async(self._h, nbytes, _ptr(density), _ptr(potential), _stream(stream)), "solve_async")

    def solve_noalloc(self, nbytes: int, density, potential, stream=None) -> None:
        """solve_async with the work area in the library's static device array
        (no allocation; FF_ERR_INVALID when the plan's work area is too large)."""
        check(lib().ff_solve_noalloc(self._h, nbytes, _ptr(density), _ptr(potential), _stream(stream)),
              "solve_noalloc")

    def stencil_apply(self, phi, out, stream=None) -> None:
        """out = L_h phi with the plan's finite-difference stencil (FF_SYMBOL_3PT /
        FF_SYMBOL_5PT plans), FP64, enqueued on stream."""
        check(lib().ff_stencil_apply(self._h, _ptr(phi), _ptr(out), _stream(stream)), "stencil_apply")

    def solve_pass(self, p: int, src, dst, stream=None) -> None:
        check(lib().ff_solve_pass(self._h, p, _ptr(src), _ptr(dst), _stream(stream)), f"pass {p}")

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(4096)
        check(lib().ff_plan_describe(self._h, buf, len(buf)), "describe")
        return buf.value.decode()

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ff_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def tune_set(kind: int, n: int, precision: int, E: int, T: int, flags: int = 0) -> None:
    """Override E / T (and FF_TUNE_SPLIT_LDS / FF_TUNE_TWO_GROUPS / FF_TUNE_LDS_TWIDDLES for column passes, FF_TUNE_GENERIC for any) for passes of
    kind (0 z forward rows, 1 y, 2 x, 3 z inverse rows) and length n."""
    t = Tune(kind, n, precision, E, T, flags)
    check(lib().ff_tune_set(ctypes.byref(t)), "tune_set")


def tune_reset() -> None:
    lib().ff_tune_reset()


def tune_query(kind: int, n: int, precision: int = 0, ncols: int = 1 << 20) -> Tune:
    t = Tune()
    check(lib().ff_tune_query(kind, n, precision, ncols, ctypes.byref(t)), "tune_query")
    return t


class Batch:
    """Many independent grids (uneven shapes, mixed precisions) launched
    together on parallel streams, optionally as one HIP graph (ff_batch_*).

    items: sequence of dicts with n, and optionally lo/hi (default: the
    benchmark grid), precision and refine_steps."""

    def __init__(self, items, flags: int = 0):
        arr = (BatchItem * len(items))()
        for k, it in enumerate(items):
            n = it["n"]
            lo, hi, _ = benchmark_grid(n)
            lo = it.get("lo", lo)
            hi = it.get("hi", hi)
            arr[k] = BatchItem(_i3(n), _d3(lo), _d3(hi), int(it.get("precision", FF_FP64)),
                               int(it.get("refine_steps", 0)))
        h = ctypes.c_void_p()
        check(lib().ff_batch_create(len(items), arr, flags, ctypes.byref(h)), "Batch")
        self._h = h
        self.count = len(items)
        self.items = list(items)

    def bytes(self, i: int) -> int:
        return int(lib().ff_batch_bytes(self._h, i))

    @staticmethod
    def _ptrs(bufs):
        return (ctypes.c_void_p * len(bufs))(*[_ptr(b) for b in bufs])

    def solve(self, densities, potentials, stream=None) -> None:
        check(lib().ff_batch_solve(self._h, self._ptrs(densities), self._ptrs(potentials), _stream(stream)),
              "batch solve")

    def wait(self, stream=None) -> None:
        check(lib().ff_batch_wait(self._h, _stream(stream)), "batch wait")

    def autotune(self, densities, potentials, reps: int = 3, stream=None):
        """Tune every distinct pass of the batch; returns (passes tuned, ms before, ms after)."""
        t = ctypes.c_int()
        b = ctypes.c_double()
        a = ctypes.c_double()
        check(lib().ff_batch_autotune(self._h, self._ptrs(densities), self._ptrs(potentials), _stream(stream), reps,
                                      ctypes.byref(t), ctypes.byref(b), ctypes.byref(a)), "batch autotune")
        return t.value, b.value, a.value

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(1024 * (self.count + 4))
        check(lib().ff_batch_describe(self._h, buf, len(buf)), "batch describe")
        return buf.value.decode()

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ff_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop (ff_server_*): admission by HBM footprint, size-class
    groups, continuous dispatch over several streams."""

    def __init__(self, hbm_budget: int = 8 << 30, streams: int = 4, max_batch: int = 16,
                 batch_window_us: float = 200.0, flags: int = 0):
        cfg = ServerConfig(hbm_budget, streams, max_batch, batch_window_us, flags)
        h = ctypes.c_void_p()
        check(lib().ff_server_create(ctypes.byref(cfg), ctypes.byref(h)), "Server")
        self._h = h
        self._keep = {}

    @staticmethod
    def request(n, precision=FF_FP64, refine_steps=0, field: Optional[Field] = None, f_host=None, phi_host=None,
                want_norms=False) -> Request:
        r = Request()
        r.n = _i3(n)
        r.precision = precision
        r.refine_steps = refine_steps
        r.field = ctypes.cast(ctypes.pointer(field), ctypes.c_void_p) if field is not None else None
        r.f_host = f_host.ctypes.data if f_host is not None else None
        r.phi_host = phi_host.ctypes.data if phi_host is not None else None
        r.want_norms = 1 if want_norms else 0
        return r

    def footprint(self, n, precision=FF_FP64, refine_steps=0, field: Optional[Field] = None,
                  want_norms=False) -> int:
        r = self.request(n, precision, refine_steps, field, None, None, want_norms)
        return int(lib().ff_server_footprint(ctypes.byref(r)))

    def submit(self, n, precision=FF_FP64, refine_steps=0, field: Optional[Field] = None, f_host=None,
               phi_host=None, want_norms=False) -> int:
        """f from a device-side field descriptor or a host array (float64, C order)."""
        import numpy as np

        if f_host is not None:
            f_host = np.ascontiguousarray(f_host, dtype=np.float64)
        r = self.request(n, precision, refine_steps, field, f_host, phi_host, want_norms)
        rid = ctypes.c_uint64()
        check(lib().ff_server_submit(self._h, ctypes.byref(r), ctypes.byref(rid)), "submit")
        self._keep[rid.value] = (f_host, phi_host)
        return rid.value

    def wait(self, rid: int, timeout: float = 120.0) -> dict:
        out = Result()
        st = lib().ff_server_wait(self._h, rid, timeout, ctypes.byref(out))
        self._keep.pop(rid, None)
        check(st, f"request {rid}")
        return {"status": out.status, "latency_s": out.latency_s, "queue_s": out.queue_s,
                "norms": tuple(out.norms), "points": out.points}

    def stats(self) -> dict:
        s = ServingStats()
        check(lib().ff_server_stats(self._h, ctypes.byref(s)), "stats")
        return {k: getattr(s, k) for k, _ in ServingStats._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ff_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def field_default(n: Sequence[int]) -> Field:
    """The benchmark's f on the 1x1x1 grid (FishFry.hpp:108-139)."""
    f = Field()
    lib().ff_field_default(_i3(n), ctypes.byref(f))
    return f


def field_set_modes(f: Field, modes) -> Field:
    """modes: iterable of ((kx, ky, kz), a, b) -> f += a cos(k.x) + b sin(k.x)."""
    modes = list(modes)
    if len(modes) > FF_MAX_MODES:
        raise FishFryError(FF_ERR_INVALID, "too many modes")
    f.nmodes = len(modes)
    for m, (k, a, b) in enumerate(modes):
        for c in range(3):
            f.mode_k[m][c] = int(k[c])
        f.mode_a[m] = float(a)
        f.mode_b[m] = float(b)
    return f


def field_init(field: Field, f, stream=None) -> None:
    check(lib().ff_field_init(ctypes.byref(field), _ptr(f), _stream(stream)), "field_init")


def field_diff(field: Field, phi, stream=None) -> Tuple[float, float, float]:
    out = (ctypes.c_double * 3)()
    check(lib().ff_field_diff(ctypes.byref(field), _ptr(phi), out, _stream(stream)), "field_diff")
    return out[0], out[1], out[2]


def field_norms(field: Field, phi, stream=None) -> Tuple[float, float, float]:
    """Relative L1, L2, Linf of FishFry.hpp:146-159,198."""
    s = field_diff(field, None, stream)
    e = field_diff(field, phi, stream)
    return e[0] / s[0], math.sqrt(e[1] / s[1]), e[2] / s[2]


from .driver import FishFry  # noqa: E402  (needs the names above)

__all__ = [
    "ParisPencil", "FishFry", "Batch", "Server", "FF_BATCH_GRAPH", "FF_BATCH_PERSISTENT", "FF_BATCH_GROUPED", "FF_BATCH_STEPS", "FF_SERVE_LARGE_ALONE", "Field", "field_default", "field_set_modes", "field_init", "field_diff",
    "field_norms", "reference_bytes", "benchmark_grid", "FishFryError", "lib", "LIB_PATH",
    "FF_FP64", "FF_FP32_REFINE", "FF_PRESERVE_INPUT", "FF_BLOCKED_LAYOUT", "FF_SLAB_PIPELINE", "FF_SYMBOL_3PT", "FF_SYMBOL_5PT", "FF_PENCIL_LAYOUT", "FF_PLAIN_LAYOUT", "FF_TUNE_SPLIT_LDS", "FF_TUNE_TWO_GROUPS", "FF_TUNE_GENERIC", "FF_TUNE_LDS_TWIDDLES",
]
