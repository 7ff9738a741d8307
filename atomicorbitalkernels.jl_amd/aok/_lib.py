"""ctypes binding of libaok.so (the C ABI in include/aok.h).

The library is the only compute path: there is no CPU fallback.  If libaok.so is
missing or cannot be loaded, every entry point raises ``AOKLibraryError``.
Device memory comes from torch tensors (``data_ptr()``); the library runs on the
context's own HIP stream and every call is blocking, like the reference's
wrappers (which end in ``KA.synchronize``).
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libaok.so")

AOK_OK = 0
AOK_ERR_ARG = 1
AOK_ERR_SPECIES = 2
AOK_ERR_LMAX = 3
AOK_ERR_DEVICE = 4
AOK_ERR_NOMEM = 5
AOK_ERR_UNSUPPORTED = 6

AOK_F32 = 0
AOK_F64 = 1
AOK_I64 = 0
AOK_I32 = 1
AOK_DECAY_GAUSSIAN = 0
AOK_DECAY_SLATER = 1
AOK_FAST = 0
AOK_PLAIN = 1
AOK_PARAMS_COEF = 0
AOK_PARAMS_PS = 1

AO_LMAX = 4
OVL_LMAX = 4


class AOKError(RuntimeError):
    """Base class of errors reported by libaok."""


class AOKLibraryError(AOKError):
    """libaok.so is not built or cannot be loaded (no fallback exists)."""


class AOKArgumentError(AOKError, ValueError):
    """Invalid argument or shape mismatch (the reference raises ErrorException /
    AssertionError before launching)."""


class AOKSpeciesError(AOKError, ValueError):
    """A species index outside 1..NZ (the reference: "species ... not in basis species list")."""


class AOKLmaxError(AOKError, ValueError):
    """Angular momentum beyond what the kernels are built for."""


class AOKDeviceError(AOKError):
    """HIP runtime failure."""


_ERRORS = {
    AOK_ERR_ARG: AOKArgumentError,
    AOK_ERR_SPECIES: AOKSpeciesError,
    AOK_ERR_LMAX: AOKLmaxError,
    AOK_ERR_DEVICE: AOKDeviceError,
    AOK_ERR_NOMEM: AOKDeviceError,
    AOK_ERR_UNSUPPORTED: AOKArgumentError,
}

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i64p = ctypes.POINTER(ctypes.c_int64)
_int = ctypes.c_int
_uint = ctypes.c_uint
_dbl = ctypes.c_double
c_void_p = ctypes.c_void_p
byref = ctypes.byref

# name -> argtypes (restype aok_status = int unless noted)
_SIGS = {
    "aok_create": [ctypes.POINTER(_vp), _int, _vp],
    "aok_destroy": [_vp],
    "aok_set_timing": [_vp, _int],
    "aok_set_blocking": [_vp, _int],
    "aok_sync": [_vp],
    "aok_kernel_times": [_vp, _i64, ctypes.POINTER(_dbl), ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(_i64)],
    "aok_tune": [_vp, ctypes.c_char_p, _i64],
    "aok_tune_get": [_vp, ctypes.c_char_p, ctypes.POINTER(_i64)],
    "aok_ao_eval": [_vp, _int, _vp, _vp, _int, _i64, _vp, _vp, _i64, _i64, _i64, _i64p, _i64p, _i64p, _i64,
                    _int, _dbl, _int, _vp, _vp],
    "aok_ao_pullback": [_vp, _int, _vp, _vp, _int, _i64, _vp, _vp, _i64, _i64, _i64, _i64p, _i64p, _i64p, _i64,
                        _int, _dbl, _int, _vp, _vp, _vp, _vp],
    "aok_compile_coef": [_vp, _int, _vp, _vp, _i64, _i64, _i64, _i64p, _vp],
    "aok_compile_coef_pb": [_vp, _int, _vp, _vp, _vp, _i64, _i64, _i64, _i64p, _vp, _vp],
    "aok_ovl2c": [_vp, _int, _uint, _vp, _vp, _vp, _vp, _int, _i64, _i64, _i64, _i64, _i64p, _vp, _vp, _dbl, _vp],
    "aok_ovl2c_noalloc": [_vp, _int, _uint, _vp, _vp, _vp, _vp, _int, _i64, _i64, _i64, _i64, _i64p, _vp, _vp, _dbl,
                          _vp],
    "aok_ovl2c_ps": [_vp, _int, _uint, _vp, _vp, _vp, _vp, _int, _i64, _i64, _i64, _i64, _i64p, _vp, _vp, _dbl,
                     _vp],
    "aok_ovl2c_pullback": [_vp, _int, _uint, _vp, _vp, _vp, _vp, _vp, _int, _i64, _i64, _i64, _i64, _i64p, _vp, _vp,
                           _dbl, _vp, _vp],
    "aok_ovl2c_pullback_ps": [_vp, _int, _uint, _vp, _vp, _vp, _vp, _vp, _int, _i64, _i64, _i64, _i64, _i64p, _vp,
                              _vp, _dbl, _vp, _vp],
    "aok_ovl2c_pullback_x": [_vp, _int, _uint, _vp, _vp, _vp, _vp, _vp, _int, _i64, _i64, _i64, _i64, _i64p, _vp,
                             _vp, _dbl, _vp, _vp],
    "aok_ovl2c_plan_pullback_x": [_vp, _vp, _vp, _vp, _vp, _vp],
    "aok_ovl2c_plan_create": [_vp, _int, _uint, _vp, _vp, _int, _i64, _i64, _i64, _i64, _i64p, _int, _vp, _vp,
                              _dbl, ctypes.POINTER(_vp)],
    "aok_ovl2c_plan_forward": [_vp, _vp, _vp, _vp],
    "aok_ovl2c_plan_pullback": [_vp, _vp, _vp, _vp, _vp, _vp],
    "aok_ovl2c_plan_update": [_vp],
    "aok_ovl2c_plan_destroy": [_vp],
    "aok_graph_begin": [_vp],
    "aok_graph_end": [_vp, ctypes.POINTER(_vp)],
    "aok_graph_launch": [_vp],
    "aok_graph_destroy": [_vp],
    "aok_ovl3c_pullback": [_vp, _int, _uint, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _int, _i64, _i64, _i64, _i64, _i64p,
                           _vp, _vp, _dbl, _vp, _vp],
    "aok_ovl3c": [_vp, _int, _uint, _vp, _vp, _vp, _vp, _vp, _vp, _int, _i64, _i64, _i64, _i64, _i64p, _vp, _vp,
                  _dbl, _vp],
}

# Every symbol include/aok.h declares (checked by the CPU test suite).
EXPORTED = sorted(list(_SIGS) + ["aok_last_error", "aok_stream", "aok_version", "aok_last_kernel_ms",
                                 "aok_last_kernel_name", "aok_get_blocking"])

_lib = None
_lock = threading.Lock()


def load():
    """Load libaok.so (once) and declare its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise AOKLibraryError(f"{LIB_PATH} is not built; run __graft_entry__.build() (no CPU fallback exists)")
        try:
            L = ctypes.CDLL(LIB_PATH)
        except OSError as e:  # pragma: no cover - depends on the box
            raise AOKLibraryError(f"cannot load {LIB_PATH}: {e}") from e
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _int
        L.aok_last_error.argtypes = [_vp]
        L.aok_last_error.restype = ctypes.c_char_p
        L.aok_stream.argtypes = [_vp]
        L.aok_stream.restype = _vp
        L.aok_version.argtypes = []
        L.aok_version.restype = ctypes.c_char_p
        L.aok_last_kernel_ms.argtypes = [_vp]
        L.aok_last_kernel_ms.restype = _dbl
        L.aok_last_kernel_name.argtypes = [_vp]
        L.aok_last_kernel_name.restype = ctypes.c_char_p
        L.aok_get_blocking.argtypes = [_vp]
        L.aok_get_blocking.restype = _int
        _lib = L
    return _lib


class Context:
    """One libaok context: its HIP stream (by default torch's default stream of the
    device, or a blocking stream ordered with it when that is the NULL stream; or the
    stream handle given), workspace and error state."""

    def __init__(self, device: int = 0, stream: int | None = None, torch_stream=None):
        self.lib = load()
        # a torch.cuda.Stream the context runs on (side_stream): calls wait for torch's
        # current stream through events instead of relying on NULL-stream ordering
        self.torch_stream = torch_stream
        self.capturing = False
        if torch_stream is not None:
            stream = torch_stream.cuda_stream
        if stream is None:
            # torch's default stream: shared when it is a real stream; when it is the legacy
            # NULL stream the library creates a blocking stream, which is ordered with it
            import torch

            stream = torch.cuda.default_stream(device).cuda_stream
        h = _vp()
        rc = self.lib.aok_create(ctypes.byref(h), int(device), _vp(stream) if stream else None)
        if rc != AOK_OK:
            raise AOKDeviceError(f"aok_create failed with status {rc}")
        self.handle = h
        self.device = device

    def close(self):
        if getattr(self, "handle", None):
            self.lib.aok_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int, what: str):
        if rc == AOK_OK:
            return
        msg = self.lib.aok_last_error(self.handle).decode(errors="replace")
        raise _ERRORS.get(rc, AOKError)(f"{what}: {msg} (status {rc})")

    @classmethod
    def side_stream(cls, device: int = 0):
        """A context on a new torch side stream (non-blocking, so it can be captured into
        a HIP graph while torch keeps the NULL stream)."""
        import torch

        return cls(device, torch_stream=torch.cuda.Stream(device))

    def call(self, name: str, *args):
        _order_inputs(self)
        rc = getattr(self.lib, name)(self.handle, *args)
        self.check(rc, name)

    @property
    def stream(self) -> int:
        return self.lib.aok_stream(self.handle)

    def set_timing(self, on: bool):
        self.lib.aok_set_timing(self.handle, 1 if on else 0)

    @property
    def blocking(self) -> bool:
        return self.lib.aok_get_blocking(self.handle) == 1

    def set_blocking(self, on: bool):
        """Blocking (default, the reference's semantics) or stream-ordered execution
        (aok_set_blocking): in stream-ordered mode calls return once enqueued on the
        context stream, torch's default stream, and device-detected errors surface at
        ``sync()``."""
        self.check(self.lib.aok_set_blocking(self.handle, 1 if on else 0), "aok_set_blocking")

    def sync(self):
        """Wait for this context's work and raise any latched device-side error."""
        self.check(self.lib.aok_sync(self.handle), "aok_sync")

    def kernel_times(self, cap: int = 65536):
        """[(kernel name, ms), ...] of the timed launches not yet returned, oldest first."""
        ms = (_dbl * cap)()
        names = (ctypes.c_char_p * cap)()
        n = _i64(0)
        self.check(self.lib.aok_kernel_times(self.handle, cap, ms, names, ctypes.byref(n)), "aok_kernel_times")
        return [(names[i].decode(), ms[i]) for i in range(n.value)]

    def tune(self, key: str, value: int):
        """Set a launch-shape knob (see aok_tune in include/aok.h)."""
        self.call("aok_tune", key.encode(), int(value))

    def tune_get(self, key: str) -> int:
        v = _i64(0)
        self.call("aok_tune_get", key.encode(), ctypes.byref(v))
        return int(v.value)

    @contextlib.contextmanager
    def capture(self):
        """Record this context's calls into a HIP graph (aok_graph_begin / _end) instead of
        running them:  ``with ctx.capture() as g: ...`` then ``g.replay()``.  Buffers,
        plans and tensors used inside stay referenced by the graph's device pointers: keep
        them alive (``g.keep(...)``) and update inputs in place.  Run the same calls once
        before capturing so that workspaces are sized.  Capture a context on a side stream
        (``Context.side_stream``) unless the library's own stream is used throughout."""
        g = Graph(self)
        self.check(self.lib.aok_graph_begin(self.handle), "aok_graph_begin")
        self.capturing = True
        ok = False
        try:
            yield g
            ok = True
        finally:
            self.capturing = False
            h = _vp()
            rc = self.lib.aok_graph_end(self.handle, ctypes.byref(h))
            if ok:
                self.check(rc, "aok_graph_end")
                g.handle = h

    def last_kernel_ms(self) -> float:
        return float(self.lib.aok_last_kernel_ms(self.handle))

    def last_kernel_name(self) -> str:
        return self.lib.aok_last_kernel_name(self.handle).decode()


class Graph:
    """A recorded sequence of one context's calls (aok_graph_*), replayed as one launch
    on the context stream.  ``keep`` pins the tensors and plans the recording uses."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.handle = None
        self._keep = []

    def keep(self, *objs):
        self._keep.extend(objs)

    def replay(self):
        _order_inputs(self.ctx)
        self.ctx.check(self.ctx.lib.aok_graph_launch(self.handle), "aok_graph_launch")

    def close(self):
        if getattr(self, "handle", None):
            self.ctx.lib.aok_graph_destroy(self.handle)
            self.handle = None
        self._keep = []

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


def _order_inputs(ctx: "Context"):
    """Order the context's next work after the inputs torch has queued.  A context on
    its own blocking stream is ordered after torch's default (NULL) stream by HIP; if
    torch is queueing on another stream, wait for it.  A context on a torch side stream
    waits for torch's current stream with an event, except while it is being captured
    (a capture records no dependencies outside itself: inputs are ordered at replay)."""
    import torch

    cur = torch.cuda.current_stream(ctx.device)
    if ctx.torch_stream is not None:
        if not ctx.capturing and cur != ctx.torch_stream:
            ctx.torch_stream.wait_stream(cur)
        return
    if cur != torch.cuda.default_stream(ctx.device):
        cur.synchronize()


def require_device(*tensors):
    """Every tensor handed to a kernel must live on one HIP device; libaok has no CPU path."""
    dev = None
    for t in tensors:
        if t is None:
            continue
        if not t.is_cuda:
            raise AOKArgumentError("inputs and outputs must be device (HIP) tensors; libaok has no CPU path "
                                   "(move a basis with adapt_basis / .to(device))")
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise AOKArgumentError(f"all tensors must be on one device; got {dev} and {t.device}")
    return dev


_ctx = {}


def context(device: int = 0) -> Context:
    """Process-wide default context per device (created on first use)."""
    c = _ctx.get(device)
    if c is None:
        c = _ctx[device] = Context(device)
    return c


def i64_array(values):
    """Host int64 array (kept alive by the caller) for the structural arguments."""
    vals = [int(v) for v in values]
    arr = (ctypes.c_int64 * max(1, len(vals)))(*vals)
    return arr
