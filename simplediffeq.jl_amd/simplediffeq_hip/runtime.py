"""Thin ctypes layer over libsimplediffeq_hip.

``Context`` owns one ``sdeh_ctx``; ``EnsembleBuffers`` lays out the
structure-of-arrays inputs/outputs of one ensemble solve (numpy arrays on the
host, or torch tensors on the GPU) and fills the C ``sdeh_io`` struct.
The same buffer object is accepted by the CPU oracle's loader in tests, so
both paths see identical bytes.
"""
import ctypes

import numpy as np

from . import _abi

_NP_DTYPE = {_abi.F64: np.float64, _abi.F32: np.float32}


class SolveError(RuntimeError):
    def __init__(self, status, message):
        self.status = status
        super().__init__("%s (%d): %s" % (_abi.STATUS_NAMES.get(status, "?"), status, message))


def make_desc(alg, dtype=_abi.F64, rhs=_abi.RHS_LORENZ, n_traj=0, tspan=(0.0, 1.0), dt=None,
              abstol=None, reltol=None, saveat=None, save_everystep=False, max_saves=0,
              maxiters=0, quirks=True, schedule=_abi.SCHED_AUTO, precision=_abi.PREC_PLAIN,
              block=0, grid=0, chunk=0, lanes=0, save_layout=_abi.SAVE_POINT_MAJOR,
              span_order=0, waves=0, per_lane=0, refill=0):
    """Build an sdeh_solve_desc.  Unset dt/abstol/reltol take the reference
    defaults (0.1f0, 1f-6, 1f-3).  lanes: lanes per trajectory (0 = for
    binary64 the right-hand side's matrix-core layout, else its group layout,
    where it has one; 1 = one lane per trajectory).  save_layout: the layout of
    u_saveat/u_steps/t_steps (point- or trajectory-major, or packed: save_everystep
    rows of all trajectories back to back, no per-trajectory cap).  waves: the
    register-cap variant (0 = tuned or default; 4, 6, 8 waves per SIMD).  per_lane:
    trajectories per lane (0 = tuned or 1; 2).  refill: the persistent schedules'
    refill batch (0 = tuned or default).  Keeps the saveat array alive on the struct."""
    d = _abi.default_desc()
    d.alg = int(alg)
    d.dtype = int(dtype)
    d.rhs = int(rhs)
    d.n_traj = int(n_traj)
    d.t0, d.tf = float(tspan[0]), float(tspan[1])
    if dt is not None:
        d.dt = float(dt)
    if abstol is not None:
        d.abstol = float(abstol)
    if reltol is not None:
        d.reltol = float(reltol)
    d.save_everystep = 1 if save_everystep else 0
    d.max_saves = int(max_saves)
    d.maxiters = int(maxiters)
    d.quirks = 1 if quirks else 0
    d.schedule = int(schedule)
    d.precision = int(precision)
    d.block, d.grid, d.chunk = int(block), int(grid), int(chunk)
    d.lanes = int(lanes)
    d.save_layout = int(save_layout)
    d.span_order = int(span_order)
    d.waves = int(waves)
    d.per_lane = int(per_lane)
    d.refill = int(refill)
    if saveat is not None and len(saveat) > 0:
        arr = np.ascontiguousarray(np.asarray(saveat, dtype=np.float64))
        d._saveat_keep = arr
        d.saveat = arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        d.n_saveat = arr.size
    return d


class EnsembleBuffers:
    """Inputs and outputs of one ensemble solve, structure-of-arrays.

    u0: (n_state, n_traj); p: (n_param, n_traj); p_shared: (n_shared,);
    tspan: (2, n_traj) or None; dt0: (n_traj,) or None; order: (n_traj,) int64 or None.
    Backend "numpy" (host) or "torch" (device tensors on `device`).
    steps_capacity: rows of the packed save_everystep output (packed layout only):
    u_steps (steps_capacity, n_state), t_steps (steps_capacity,), and
    steps_offset (n_traj + 1,) int64; trajectory i owns rows
    steps_offset[i] .. steps_offset[i + 1] - 1 (see steps_of)."""

    def __init__(self, desc, n_state, u0, p=None, p_shared=None, tspan=None, dt0=None,
                 order=None, backend="numpy", device=None, steps_capacity=0):
        self.desc = desc
        self.backend = backend
        n = int(desc.n_traj)
        dt_np = _NP_DTYPE[desc.dtype]
        self.n_state = n_state
        if backend == "numpy":
            conv = lambda a, dt: None if a is None else np.ascontiguousarray(np.asarray(a, dtype=dt))
            empty = lambda shape, dt: np.empty(shape, dtype=dt)
            ptr = lambda a: None if a is None else a.ctypes.data
        elif backend == "torch":
            import torch
            tdt = {np.float64: torch.float64, np.float32: torch.float32, np.int32: torch.int32,
                   np.int64: torch.int64}

            def conv(a, dt):
                if a is None:
                    return None
                if isinstance(a, torch.Tensor):
                    return a.to(device=device, dtype=tdt[dt]).contiguous()
                return torch.as_tensor(np.asarray(a, dtype=dt), device=device).contiguous()

            empty = lambda shape, dt: torch.empty(shape, dtype=tdt[dt], device=device)
            ptr = lambda a: None if a is None else a.data_ptr()
        else:
            raise ValueError("backend must be 'numpy' or 'torch'")
        self.u0 = conv(u0, dt_np)
        self.p = conv(p, dt_np)
        self.p_shared = conv(p_shared, dt_np)
        self.tspan = conv(tspan, np.float64)
        self.dt0 = conv(dt0, np.float64)
        self.order = conv(order, np.int64)
        self.u_end = empty((n_state, n), dt_np)
        self.t_end = empty((n,), np.float64)
        self.retcode = empty((n,), np.int32)
        self.naccept = empty((n,), np.int32)
        self.nreject = empty((n,), np.int32)
        ns = int(desc.n_saveat)
        ms = int(desc.max_saves)
        tm = int(desc.save_layout) == _abi.SAVE_TRAJ_MAJOR
        self.traj_major = tm
        self.u_saveat = None
        if ns > 0:
            self.u_saveat = empty((n, ns, n_state) if tm else (ns, n_state, n), dt_np)
        everystep = ns == 0 and desc.save_everystep
        self.packed = int(desc.save_layout) == _abi.SAVE_PACKED
        self.u_steps = self.t_steps = self.steps_offset = None
        self.steps_capacity = int(steps_capacity)
        if everystep and self.packed:
            cap = self.steps_capacity
            self.u_steps = empty((cap, n_state), dt_np)
            self.t_steps = empty((cap,), np.float64)
            self.steps_offset = empty((n + 1,), np.int64)
        elif everystep and ms > 0:
            self.u_steps = empty((n, ms, n_state) if tm else (ms, n_state, n), dt_np)
            self.t_steps = empty((n, ms) if tm else (ms, n), np.float64)
        self.n_steps = empty((n,), np.int32) if everystep else None
        io = _abi.SolveIO()
        io.u0 = ptr(self.u0)
        io.p = ptr(self.p)
        io.p_shared = ptr(self.p_shared)
        io.n_shared = 0 if self.p_shared is None else int(np.prod(tuple(self.p_shared.shape)))
        io.tspan = ptr(self.tspan)
        io.dt0 = ptr(self.dt0)
        io.order = ptr(self.order)
        io.u_end = ptr(self.u_end)
        io.t_end = ptr(self.t_end)
        io.retcode = ptr(self.retcode)
        io.naccept = ptr(self.naccept)
        io.nreject = ptr(self.nreject)
        io.u_saveat = ptr(self.u_saveat)
        io.u_steps = ptr(self.u_steps)
        io.t_steps = ptr(self.t_steps)
        io.n_steps = ptr(self.n_steps)
        io.steps_offset = ptr(self.steps_offset)
        io.steps_capacity = self.steps_capacity if self.packed else 0
        self.io = io
        # Device buffers are filled (copies, dtype conversions) on torch's current
        # stream; a solve on another stream waits for this event first.
        self.ready = None
        if backend == "torch":
            self.ready = torch.cuda.Event()
            self.ready.record(torch.cuda.current_stream(self.u0.device))

    def steps_of(self, i):
        """(t, u) of trajectory i's save_everystep rows in the packed layout, as
        numpy arrays (t: (k,), u: (k, n_state))."""
        off = self.numpy("steps_offset")
        a, b = int(off[i]), int(off[i + 1])
        if self.backend == "numpy":
            return self.t_steps[a:b], self.u_steps[a:b]
        return self.t_steps[a:b].cpu().numpy(), self.u_steps[a:b].cpu().numpy()

    def numpy(self, name):
        a = getattr(self, name)
        if a is None or self.backend == "numpy":
            return a
        return a.cpu().numpy()


class Context:
    """One library context on a HIP device."""

    def __init__(self, device=0):
        self.lib = _abi.load_library()
        h = ctypes.c_void_p()
        st = self.lib.sdeh_create(ctypes.byref(h), int(device))
        if st != 0:
            raise SolveError(st, "sdeh_create(device=%d) failed" % device)
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self.lib.sdeh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, allow=()):
        if st != 0 and st not in allow:
            raise SolveError(st, self.lib.sdeh_last_error(self.h).decode())
        return st

    def last_error(self):
        return self.lib.sdeh_last_error(self.h).decode()

    def rhs_info(self, rhs):
        n, p, s = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        self._check(self.lib.sdeh_rhs_info(self.h, int(rhs), ctypes.byref(n), ctypes.byref(p),
                                           ctypes.byref(s)))
        return n.value, p.value, s.value

    def compile_rhs(self, src, fn_name, n_state, n_param, n_shared=0):
        h = ctypes.c_int64()
        self._check(self.lib.sdeh_compile_rhs(self.h, src.encode(), fn_name.encode(),
                                              int(n_state), int(n_param), int(n_shared),
                                              ctypes.byref(h)))
        return h.value

    @staticmethod
    def _torch_stream(bufs):
        """The stream for a solve on device buffers: torch's current stream, so
        that work queued there (the kernels that produced the inputs) completes
        before the solve starts, as for any torch operation.  torch's default
        stream has handle 0, which the library reads as its own stream: then
        torch's stream is synchronised first and None is returned."""
        import torch
        ts = torch.cuda.current_stream(bufs.u0.device)
        if ts.cuda_stream:
            return ts.cuda_stream
        ts.synchronize()
        return None

    @staticmethod
    def _after_buffers(bufs, stream):
        """Order a solve on the explicit `stream` after the work that filled the device
        buffers on torch's stream (without it the kernels could read inputs that a
        conversion on the other stream is still writing)."""
        if bufs.backend == "torch" and stream and bufs.ready is not None:
            import torch
            torch.cuda.ExternalStream(stream, device=bufs.u0.device).wait_event(bufs.ready)

    def solve(self, desc, bufs, stream=None, sync=True, allow=(_abi.E_DTMIN,)):
        """Run sdeh_solve on an EnsembleBuffers; returns the status.  Device
        (torch) buffers run on torch's current stream unless `stream` is given; a
        given stream first waits for the buffers' construction."""
        mem = _abi.MEM_HOST if bufs.backend == "numpy" else _abi.MEM_DEVICE
        if stream is None and bufs.backend == "torch":
            stream = self._torch_stream(bufs)
        else:
            self._after_buffers(bufs, stream)
        st = self.lib.sdeh_solve(self.h, ctypes.byref(desc), ctypes.byref(bufs.io), mem,
                                 ctypes.c_void_p(stream), 1 if sync else 0)
        return self._check(st, allow)

    def solve_static(self, desc, bufs, stream=None, sync=True, allow=(_abi.E_DTMIN,)):
        """sdeh_solve_static (no device allocation) on device (torch) EnsembleBuffers;
        returns the status."""
        if bufs.backend != "torch":
            raise ValueError("the allocation-free solve takes device buffers")
        if stream is None:
            stream = self._torch_stream(bufs)
        else:
            self._after_buffers(bufs, stream)
        st = self.lib.sdeh_solve_static(self.h, ctypes.byref(desc), ctypes.byref(bufs.io),
                                        ctypes.c_void_p(stream), 1 if sync else 0)
        return self._check(st, allow)

    def stats(self):
        s = _abi.Stats()
        self._check(self.lib.sdeh_get_stats(self.h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _abi.Stats._fields_}

    def autotune(self, desc, bufs):
        if bufs.backend == "torch":  # the tuner runs on the context's stream
            import torch
            torch.cuda.current_stream(bufs.u0.device).synchronize()
        s = _abi.Stats()
        self._check(self.lib.sdeh_autotune(self.h, ctypes.byref(desc), ctypes.byref(bufs.io),
                                           ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in _abi.Stats._fields_}

    def graph_capture(self, desc, bufs):
        if bufs.backend == "torch":  # capture runs on the context's stream
            import torch
            torch.cuda.current_stream(bufs.u0.device).synchronize()
        h = ctypes.c_int64()
        self._check(self.lib.sdeh_graph_capture(self.h, ctypes.byref(desc), ctypes.byref(bufs.io),
                                                ctypes.byref(h)))
        return h.value

    def graph_launch(self, handle, stream=None):
        self._check(self.lib.sdeh_graph_launch(self.h, int(handle), ctypes.c_void_p(stream)))

    def graph_destroy(self, handle):
        self._check(self.lib.sdeh_graph_destroy(self.h, int(handle)))

    def synchronize(self):
        self._check(self.lib.sdeh_synchronize(self.h))

    # MLP cost model (device pointers, e.g. torch tensor data_ptr()).
    def mlp_forward(self, params, x, y, n, stream=None):
        self._check(self.lib.sdeh_mlp_forward(self.h, params, x, y, int(n), ctypes.c_void_p(stream)))

    def mlp_plan_lorenz(self, params, p, u0, tspan, n, feat, y_norm, key, dt0, pred=None,
                        stream=None):
        """sdeh_mlp_plan_lorenz on device pointers; feat: 8 float64, y_norm: 4 float32
        (host sequences, kept alive for the call)."""
        fa = (ctypes.c_double * 8)(*[float(v) for v in feat])
        ya = (ctypes.c_float * 4)(*[float(v) for v in y_norm])
        self._check(self.lib.sdeh_mlp_plan_lorenz(
            self.h, params, p, u0, tspan, int(n), ctypes.cast(fa, ctypes.c_void_p),
            ctypes.cast(ya, ctypes.c_void_p), key, dt0, pred, ctypes.c_void_p(stream)))

    def plan_order(self, key, n, order, stream=None):
        """sdeh_plan_order on device pointers: order (int64 [n]) by increasing FP32
        key (decreasing predicted cost), bucketed at 1/16 octave of exp(-key)."""
        self._check(self.lib.sdeh_plan_order(self.h, key, int(n), order,
                                             ctypes.c_void_p(stream)))

    def mlp_loss_grad(self, params, x, y, n, loss, grad, stream=None):
        self._check(self.lib.sdeh_mlp_loss_grad(self.h, params, x, y, int(n), loss, grad,
                                                ctypes.c_void_p(stream)))

    def mlp_adam(self, params, grad, m, v, lr, beta1, beta2, eps, t, stream=None):
        self._check(self.lib.sdeh_mlp_adam(self.h, params, grad, m, v, float(lr), float(beta1),
                                           float(beta2), float(eps), int(t),
                                           ctypes.c_void_p(stream)))


def fixed_grid(t0, dt, tf, cap=0):
    """length(t0:dt:tf) and its first `cap` elements (Julia Float64 range rules)."""
    lib = _abi.load_library()
    lib.sdeh_fixed_grid.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_void_p, ctypes.c_int64]
    lib.sdeh_fixed_grid.restype = ctypes.c_int64
    out = np.empty(max(cap, 1), dtype=np.float64)
    n = lib.sdeh_fixed_grid(float(t0), float(dt), float(tf), out.ctypes.data, int(cap))
    return n, out[:min(cap, max(n, 0))]
