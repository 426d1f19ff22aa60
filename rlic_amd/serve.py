"""Serving: continuous batching of ``convolve`` requests on one GPU.

A :class:`Server` accepts ``convolve`` requests at any time (:meth:`Server.submit`).
It admits them into device memory, first come first served, as its byte
budget allows, and :meth:`Server.step` advances every admitted request by
one pass (iteration), with all admitted requests of a dtype in one
persistent launch (``rlic_server_*`` of ``librlic_hip.so``).  Requests that
arrive while others run join at the next step, and their passes run
together with the others' on the device.  Whatever else shares the steps,
a request's result is bit-identical to ``rlic_amd.convolve`` with the same
arguments.

With ``keep=True`` a request's field (its packed vectors, move record and
last output) stays resident after the request finishes, and
:meth:`Server.follow_up` continues it with more passes from its last
output: the recorded moves are replayed, no streamline is traced again,
and the result equals ``convolve`` with the total number of iterations.
Kept fields are freed by :meth:`Server.release` or, when an admission
needs their memory, least recently used first; a follow-up of an evicted
field is admitted again from its original inputs (same result, more work).

The caller may drive the steps itself (:meth:`Server.step`,
:meth:`Server.run`) or let a background thread do it (:meth:`Server.start`
... :meth:`Server.stop`): submissions from any thread are then admitted at
the next step boundary while the current step runs on the GPU (the native
step releases the GIL), and :meth:`Request.wait` blocks until a request is
done.
"""

from __future__ import annotations

__all__ = ["Request", "Server"]

import ctypes
import itertools
import threading
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from rlic_amd import _core, _native
from rlic_amd._lib import validate

_MODE = {"velocity": 0, "polarization": 1}
_BOUND = {"closed": 0, "periodic": 1}
_ids = itertools.count(1)


@dataclass(eq=False)
class Request:
    """One ``convolve`` call, or a follow-up of a kept one."""

    texture: Any
    u: Any
    v: Any
    kernel: Any
    uv_mode: str
    boundaries: Any
    iterations: int            # passes this request asks for
    keep: bool
    arrival: float             # clock time the request arrived
    base: Request | None = None  # the kept request this one follows up
    id: int = field(default_factory=lambda: next(_ids))
    state: str = "queued"      # queued -> active -> done | failed
    admitted: float | None = None
    finished: float | None = None
    result: Any = None
    error: BaseException | None = None
    total_iterations: int = 0  # passes of the field after this request
    pixel_steps: int = 0       # work of this request (H*W*(L-1)*iterations)
    _field: _Field | None = None
    _event: threading.Event = field(default_factory=threading.Event, repr=False)

    @property
    def latency(self) -> float | None:
        return None if self.finished is None else self.finished - self.arrival

    def wait(self, timeout: float | None = None) -> bool:
        """Block until the request is done or failed; False on timeout."""
        return self._event.wait(timeout)

    def _finish(self, state: str, when: float) -> None:
        self.state = state
        self.finished = when
        self._event.set()


@dataclass(eq=False)
class _Field:
    """A field resident in the native server."""

    handle: int
    dtype: np.dtype
    shape: tuple[int, int]
    keep: bool
    passes: int                # passes requested so far (cumulative)
    request: Request | None    # the request its running passes belong to
    last_used: float
    inputs: tuple              # (texture, u, v, kernel) device tensors, for re-admission
    boundaries: Any
    uv_mode: str


class Server:
    """Continuous-batching LIC server on one device (see the module docstring).

    submit, follow_up and release may be called from any thread; steps run
    one at a time (on the background thread after start(), else on the
    caller's).
    """

    def __init__(self, budget_bytes: int = 64 << 30, *, device=None, profile=None,
                 clock=time.perf_counter):
        import torch

        self._torch = torch
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.profile = _core.DEFAULT_PROFILE if profile is None else (
            _core.ARITHMETIC_PROFILES[profile] if isinstance(profile, str) else int(profile))
        self.budget_bytes = int(budget_bytes)
        self.clock = clock
        self._lib = _native.load()
        with torch.cuda.device(self.device):
            self._stream = torch.cuda.current_stream(self.device)
            ptr = self._lib.rlic_server_create(self.budget_bytes, self.profile,
                                               ctypes.c_void_p(self._stream.cuda_stream))
        if not ptr:
            _native.raise_for_status(_native.RLIC_ERR_INVALID_ARGUMENT)
        self._srv = ctypes.c_void_p(ptr)
        self._lock = threading.RLock()            # Python-side state
        self._cv = threading.Condition(self._lock)
        self._step_lock = threading.Lock()        # one step at a time
        self._thread: threading.Thread | None = None
        self._stopping = False
        self.loop_error: BaseException | None = None
        self._queue: deque[Request] = deque()
        self._active: dict[int, _Field] = {}
        self._idle: OrderedDict[int, _Field] = OrderedDict()  # kept, finished; LRU first
        self.steps = 0
        self.evictions = 0
        self.readmissions = 0

    # ------------------------------------------------------------------ API
    def submit(self, texture, u, v, *, kernel, uv_mode="velocity", boundaries="closed",
               iterations: int = 1, keep: bool = False, arrival: float | None = None) -> Request:
        """Queue one ``convolve`` call; validates it as ``convolve`` does."""
        bs, dtype = validate(texture, u, v, kernel=kernel, uv_mode=uv_mode,
                             boundaries=boundaries, iterations=iterations)
        now = self.clock() if arrival is None else arrival
        req = Request(texture, u, v, kernel, uv_mode, (bs.x, bs.y), int(iterations), bool(keep),
                      now)
        req.total_iterations = req.iterations
        ny, nx = texture.shape
        req.pixel_steps = int(ny) * int(nx) * max(int(kernel.shape[0]) - 1, 0) * req.iterations
        if req.iterations == 0:  # src/rlic/_lib.py:208-209: a copy of the texture
            req.result = texture.clone() if _core._is_torch(texture) else texture.copy()
            req.admitted = self.clock()
            req._finish("done", req.admitted)
            return req
        with self._cv:
            self._queue.append(req)
            self._cv.notify_all()
        return req

    def follow_up(self, request: Request, iterations: int, *,
                  arrival: float | None = None) -> Request:
        """Queue ``iterations`` more passes of a finished ``keep=True`` request."""
        if not request.keep:
            raise ValueError("only a request submitted with keep=True can be followed up")
        if request.state != "done":
            raise ValueError("the request to follow up has not finished")
        if iterations < 0:
            raise ValueError(f"Invalid number of iterations: {iterations}\n"
                             "Expected a strictly positive integer.")
        now = self.clock() if arrival is None else arrival
        req = Request(request.texture, request.u, request.v, request.kernel, request.uv_mode,
                      request.boundaries, int(iterations), True, now, base=request)
        req.total_iterations = request.total_iterations + req.iterations
        ny, nx = request.texture.shape
        req.pixel_steps = int(ny) * int(nx) * max(int(request.kernel.shape[0]) - 1, 0) * \
            req.iterations
        if req.iterations == 0:
            req.result = request.result
            req._field = request._field
            req.admitted = self.clock()
            req._finish("done", req.admitted)
            return req
        with self._cv:
            self._queue.append(req)
            self._cv.notify_all()
        return req

    def step(self) -> list[Request]:
        """Admit what fits, run one pass of every admitted request, and
        return the requests that finished in this step."""
        with self._step_lock:
            with self._lock:
                self._admit()
                if not self._active:
                    return []
                cap = len(self._active)
            out = (ctypes.c_int64 * cap)()
            n = ctypes.c_int64(0)
            # the GPU work of the step; other threads may submit meanwhile
            _native.raise_for_status(
                self._lib.rlic_server_step(self._srv, out, cap, ctypes.byref(n)))
            with self._cv:
                self.steps += 1
                done = []
                now = self.clock()
                for h in out[: n.value]:
                    fld = self._active.pop(h)
                    req = fld.request
                    req.result = self._read(fld)
                    fld.request = None
                    fld.last_used = now
                    if fld.keep:
                        self._idle[h] = fld
                    else:
                        self._release_handle(h)
                    req._finish("done", self.clock())
                    done.append(req)
                self._cv.notify_all()
            return done

    def run(self) -> None:
        """Until no request is queued or running: step (or, with the
        background thread running, wait for it)."""
        if self._thread is not None:
            with self._cv:
                while (self._queue or self._active) and self.loop_error is None:
                    self._cv.wait(0.05)
            if self.loop_error is not None:
                raise self.loop_error
            return
        while self._queue or self._active:
            before = len(self._queue) + len(self._active)
            self.step()
            if not self._active and self._queue and len(self._queue) == before:
                # nothing could be admitted and nothing runs: the head cannot fit
                raise MemoryError("a queued request does not fit the serving budget")

    def start(self) -> None:
        """Run steps on a background thread whenever requests are pending."""
        if self._thread is not None:
            return
        self._stopping = False
        self._thread = threading.Thread(target=self._loop, name="rlic-serve", daemon=True)
        self._thread.start()

    def stop(self) -> None:
        """Finish the pending requests, then end the background thread."""
        th = self._thread
        if th is None:
            return
        with self._cv:
            self._stopping = True
            self._cv.notify_all()
        th.join()
        self._thread = None

    def _loop(self) -> None:
        while True:
            with self._cv:
                while not (self._queue or self._active):
                    if self._stopping:
                        return
                    self._cv.wait()
            try:
                before = (len(self._queue), len(self._active))
                self.step()
            except BaseException as exc:  # reported to waiters
                with self._cv:
                    self.loop_error = exc
                    for req in list(self._queue):
                        req.error = exc
                        req._finish("failed", self.clock())
                    self._queue.clear()
                    for fld in self._active.values():
                        if fld.request is not None:
                            fld.request.error = exc
                            fld.request._finish("failed", self.clock())
                    self._active.clear()
                    self._cv.notify_all()
                return
            with self._cv:
                if not self._active and (len(self._queue), len(self._active)) == before:
                    self._cv.wait(0.01)  # nothing admissible yet: wait for a change

    def release(self, request: Request) -> None:
        """Free a finished kept request's resident field."""
        with self._lock:
            fld = request._field
            if fld is not None and fld.handle in self._idle:
                del self._idle[fld.handle]
                self._release_handle(fld.handle)

    def close(self) -> None:
        if getattr(self, "_thread", None) is not None:
            self.stop()
        if getattr(self, "_srv", None):
            self._lib.rlic_server_destroy(self._srv)
            self._srv = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def is_resident(self, request: Request) -> bool:
        """Whether the request's field is in device memory (running or kept)."""
        fld = request._field
        return fld is not None and (fld.handle in self._idle or fld.handle in self._active)

    @property
    def queued(self) -> int:
        return len(self._queue)

    @property
    def active(self) -> int:
        return len(self._active)

    def usage(self) -> dict:
        used, resident, active = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._lib.rlic_server_usage(self._srv, ctypes.byref(used), ctypes.byref(resident),
                                    ctypes.byref(active))
        return {"used_bytes": used.value, "resident": resident.value, "active": active.value,
                "queued": len(self._queue), "budget_bytes": self.budget_bytes}

    # ------------------------------------------------------------ internals
    def _field_bytes(self, dtype, ny, nx, klen, keep, iterations) -> int:
        fn = (self._lib.rlic_server_field_bytes_f32 if dtype == np.float32
              else self._lib.rlic_server_field_bytes_f64)
        return int(fn(ny, nx, klen, int(keep), iterations))

    def _free_bytes(self) -> int:
        return self.budget_bytes - self.usage()["used_bytes"]

    def _make_room(self, need: int, protect: _Field | None = None) -> bool:
        """Evict idle kept fields, least recently used first, until `need`
        bytes are free; False if that is impossible."""
        free = self._free_bytes()
        while free < need:
            victim = next((h for h, f in self._idle.items() if f is not protect), None)
            if victim is None:
                return False
            fld = self._idle.pop(victim)
            self._release_handle(victim)
            self.evictions += 1
            free = self._free_bytes()
        return True

    def _admit(self) -> None:
        torch = self._torch
        blocked_new = False
        for req in list(self._queue):
            base = req.base._field if req.base is not None else None
            if req.base is not None and base is not None and base.handle in self._idle:
                # follow-up of a resident field: add passes, no new memory
                fld = self._idle.pop(base.handle)
                _native.raise_for_status(
                    self._lib.rlic_server_extend(self._srv, fld.handle, req.iterations))
                fld.passes += req.iterations
                self._start(req, fld)
                continue
            if req.base is not None and base is not None and base.handle in self._active:
                continue  # its field is still busy with an earlier follow-up
            if blocked_new:
                continue
            # a new field (or a follow-up whose field was evicted: admitted
            # again from its inputs for all its passes)
            passes = req.total_iterations if req.base is not None else req.iterations
            tensors = self._device_inputs(req)
            tex = tensors[0]
            dtype = np.float32 if tex.dtype == torch.float32 else np.float64
            ny, nx = (int(s) for s in tex.shape)
            klen = int(tensors[3].shape[0])
            need = self._field_bytes(dtype, ny, nx, klen, req.keep, passes)
            if need > self.budget_bytes:
                self._queue.remove(req)
                req.error = MemoryError(f"the field needs {need} bytes, more than the serving "
                                        f"budget of {self.budget_bytes}")
                req._finish("failed", self.clock())
                continue
            if not self._make_room(need):
                blocked_new = True  # first come, first served for new fields
                continue
            handle = self._admit_native(req, tensors, dtype, ny, nx, klen, passes)
            fld = _Field(handle, np.dtype(dtype), (ny, nx), req.keep, passes, None,
                         self.clock(), tuple(tensors), req.boundaries, req.uv_mode)
            if req.base is not None:
                self.readmissions += 1
            self._start(req, fld)

    def _start(self, req: Request, fld: _Field) -> None:
        self._queue.remove(req)
        req.state = "active"
        req.admitted = self.clock()
        req._field = fld
        fld.request = req
        self._active[fld.handle] = fld

    def _device_inputs(self, req: Request) -> list:
        torch = self._torch
        src = req.base if req.base is not None else req
        while src.base is not None:
            src = src.base
        if src._field is not None and src._field.inputs:
            return list(src._field.inputs)
        arrays = (src.texture, src.u, src.v, src.kernel)
        if _core._is_torch(arrays[0]):
            tensors = list(arrays)
        else:
            tensors = [torch.from_numpy(np.ascontiguousarray(a)).to(self.device) for a in arrays]
        tensors[3] = tensors[3].contiguous()
        return tensors

    def _admit_native(self, req, tensors, dtype, ny, nx, klen, passes) -> int:
        tex, u, v, k = tensors
        struct = _native.FieldF32 if dtype == np.float32 else _native.FieldF64
        entry = (self._lib.rlic_server_admit_f32 if dtype == np.float32
                 else self._lib.rlic_server_admit_f64)
        ctype = ctypes.c_float if dtype == np.float32 else ctypes.c_double

        def ptr(t):
            return ctypes.cast(ctypes.c_void_p(t.data_ptr()), ctypes.POINTER(ctype))

        f = struct()
        f.texture, f.tex_sy, f.tex_sx = ptr(tex), tex.stride(0), tex.stride(1)
        f.u, f.u_sy, f.u_sx = ptr(u), u.stride(0), u.stride(1)
        f.v, f.v_sy, f.v_sx = ptr(v), v.stride(0), v.stride(1)
        f.ny, f.nx = ny, nx
        f.kernel, f.kernel_len = ptr(k), klen
        (xl, xr), (yl, yr) = req.boundaries
        f.uv_mode = _MODE[req.uv_mode]
        f.x_left, f.x_right, f.y_left, f.y_right = (_BOUND[xl], _BOUND[xr], _BOUND[yl],
                                                     _BOUND[yr])
        f.iterations = passes
        f.out = ptr(tex)
        handle = ctypes.c_int64(0)
        _native.raise_for_status(entry(self._srv, ctypes.byref(f), int(req.keep),
                                       ctypes.byref(handle)))
        return handle.value

    def _read(self, fld: _Field):
        torch = self._torch
        tdtype = torch.float32 if fld.dtype == np.float32 else torch.float64
        out = torch.empty(fld.shape, dtype=tdtype, device=self.device)
        fn = self._lib.rlic_server_read_f32 if fld.dtype == np.float32 else \
            self._lib.rlic_server_read_f64
        if out.numel() > 0:
            _native.raise_for_status(fn(self._srv, fld.handle, ctypes.c_void_p(out.data_ptr())))
        req = fld.request
        src = req
        while src.base is not None:
            src = src.base
        return out if _core._is_torch(src.texture) else out.cpu().numpy()

    def _release_handle(self, handle: int) -> None:
        _native.raise_for_status(self._lib.rlic_server_release(self._srv, handle))
