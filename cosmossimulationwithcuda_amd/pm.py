"""Python binding of libpm.so (include/pm.h) over ctypes.

Mirrors the compute interface of the reference simulator (CosmosCuda.cuh,
struct Universe): create with a configuration (Constants + ctor arguments,
cuh:29-74, 881), load particle arrays, step (worker loop body, cuh:1336-1341),
read state back.  There is no CPU fallback: every call goes to the HIP library
and a missing library or device raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpm.so")

PM_OK, PM_EINVAL, PM_ENOMEM, PM_EDEVICE, PM_EBUSY, PM_ECANCELLED, PM_ENODEV = range(7)
NGP, CIC = 0, 1
WRAP_KEEP_VELOCITY, WRAP_REFERENCE = 0, 1
FORCE_SPECTRAL, FORCE_STENCIL = 0, 1
FFT_FP32, FFT_BF16X3, FFT_BF16X6 = 0, 1, 2
LATTICE_MASS, LATTICE_POTENTIAL, LATTICE_FORCE, LATTICE_TRANSFER = range(4)
STAGE_DEPOSIT, STAGE_FIELD, STAGE_PUSH = range(3)
TUNE_FFT, TUNE_PARTICLES = 1, 2
KERNELS = ["clear", "deposit", "row_r2c", "col_conv", "row_c2r", "edge", "stencil", "push", "sort", "push_deposit",
           "push_deposit_sort", "push_deposit_count"]


class PMError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"pm status {status}: {msg}")
        self.status = status


class Config(ctypes.Structure):
    _fields_ = [
        ("n_lattice", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("n_particles", ctypes.c_int64),
        ("dt", ctypes.c_float),
        ("gravity", ctypes.c_float),
        ("green_origin", ctypes.c_double),
        ("deconv_eps", ctypes.c_double),
        ("wrap_mode", ctypes.c_int32),
        ("force_mode", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("mass_shift", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("resort_interval", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 6),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("n_lattice", ctypes.c_int32),
        ("mass_shift", ctypes.c_int32),
        ("n_particles", ctypes.c_int64),
        ("capacity", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
        ("device", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("sorts", ctypes.c_int64),
        ("migrants", ctypes.c_int64),
    ]


_lib = None
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_H = ctypes.c_void_p


def lib():
    """Load libpm.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m cosmossimulationwithcuda_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.pm_config_default.argtypes = [ctypes.POINTER(Config)]
        L.pm_instance_bytes.argtypes = [ctypes.POINTER(Config)]
        L.pm_instance_bytes.restype = ctypes.c_int64
        L.pm_device_count.restype = ctypes.c_int
        L.pm_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(_H)]
        L.pm_destroy.argtypes = [_H]
        L.pm_load.argtypes = [_H, _f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_int64]
        L.pm_load_verlet.argtypes = [_H, _f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_int64]
        L.pm_step.argtypes = [_H, ctypes.c_int32]
        L.pm_stage.argtypes = [_H, ctypes.c_int32]
        L.pm_sync.argtypes = [_H]
        L.pm_read.argtypes = [_H, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.pm_read_verlet.argtypes = [_H, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.pm_read_lattice.argtypes = [_H, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t]
        L.pm_diagnostics.argtypes = [_H, ctypes.POINTER(ctypes.c_double)]
        L.pm_get_info.argtypes = [_H, ctypes.POINTER(Info)]
        L.pm_frame.argtypes = [_H, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
        L.pm_checkpoint_bytes.argtypes = [_H]
        L.pm_checkpoint_bytes.restype = ctypes.c_int64
        L.pm_checkpoint_write.argtypes = [_H, ctypes.c_void_p, ctypes.c_int64]
        L.pm_checkpoint_restore.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(_H)]
        L.pm_checkpoint_save.argtypes = [_H, ctypes.c_char_p]
        L.pm_checkpoint_open.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(_H)]
        L.pm_set_param.argtypes = [_H, ctypes.c_char_p, ctypes.c_int64]
        L.pm_get_param.argtypes = [_H, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]
        L.pm_autotune.argtypes = [_H, ctypes.c_int32, ctypes.c_int32]
        L.pm_profile.argtypes = [_H, ctypes.c_int32]
        L.pm_profile_read.argtypes = [_H, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_int64)]
        L.pm_profile_reset.argtypes = [_H]
        L.pm_diagnostics_noalloc.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32] + \
            [ctypes.c_void_p] * 5 + [ctypes.c_int64, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
        L.pm_stream.argtypes = [_H]
        L.pm_stream.restype = ctypes.c_void_p
        L.pm_last_error.argtypes = [_H]
        L.pm_last_error.restype = ctypes.c_char_p
        L.pm_status_string.argtypes = [ctypes.c_int]
        L.pm_status_string.restype = ctypes.c_char_p
        _declare_server(L)
        _lib = L
    return _lib


def device_count():
    return lib().pm_device_count()


def default_config(**kw):
    c = Config()
    lib().pm_config_default(ctypes.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


def instance_bytes(cfg):
    return lib().pm_instance_bytes(ctypes.byref(cfg))


def _check(st, h=None):
    if st != PM_OK:
        msg = lib().pm_last_error(h).decode(errors="replace")
        raise PMError(st, msg)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Simulation:
    """One simulation instance on one GPU (one HIP stream)."""

    def __init__(self, cfg=None, **kw):
        self.cfg = cfg if cfg is not None else default_config(**kw)
        self.n = self.cfg.n_lattice
        h = _H()
        _check(lib().pm_create(ctypes.byref(self.cfg), ctypes.byref(h)))
        self._h = h

    # lifecycle
    def close(self):
        if getattr(self, "_h", None):
            _check(lib().pm_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @classmethod
    def _adopt(cls, h):
        self = cls.__new__(cls)
        self._h = h
        info = Info()
        _check(lib().pm_get_info(h, ctypes.byref(info)), h)
        self.cfg = default_config(n_lattice=info.n_lattice, n_particles=info.capacity, device=info.device)
        self.n = info.n_lattice
        return self

    # checkpoints
    def checkpoint(self):
        """The state as checkpoint bytes (header + float32 x, y, vx, vy, m)."""
        nb = lib().pm_checkpoint_bytes(self._h)
        buf = np.empty(nb, np.uint8)
        _check(lib().pm_checkpoint_write(self._h, buf.ctypes.data, nb), self._h)
        return buf.tobytes()

    def save(self, path):
        _check(lib().pm_checkpoint_save(self._h, os.fsencode(path)), self._h)

    # state
    def load(self, x, y, vx, vy, m):
        arrs = [_f32(a) for a in (x, y, vx, vy, m)]
        n = len(arrs[0])
        if any(len(a) != n for a in arrs):
            raise ValueError("particle arrays differ in length")
        _check(lib().pm_load(self._h, *arrs, n), self._h)

    def load_verlet(self, x, y, x_old, y_old, m):
        arrs = [_f32(a) for a in (x, y, x_old, y_old, m)]
        _check(lib().pm_load_verlet(self._h, *arrs, len(arrs[0])), self._h)

    def set_param(self, name, value):
        _check(lib().pm_set_param(self._h, name.encode(), int(value)), self._h)

    def get_param(self, name):
        v = ctypes.c_int64()
        _check(lib().pm_get_param(self._h, name.encode(), ctypes.byref(v)), self._h)
        return v.value

    def autotune(self, fft=True, particles=True, reps=5):
        flags = (TUNE_FFT if fft else 0) | (TUNE_PARTICLES if particles else 0)
        _check(lib().pm_autotune(self._h, flags, reps), self._h)
        return {k: self.get_param(k) for k in ("fft_rp", "fft_cb", "fft_radix", "fft_it", "tile_log", "deposit_chunk",
                                                 "resort_interval")}

    def info(self):
        i = Info()
        _check(lib().pm_get_info(self._h, ctypes.byref(i)), self._h)
        return i

    def read(self):
        n = self.info().n_particles
        out = [np.empty(n, np.float32) for _ in range(4)]
        _check(lib().pm_read(self._h, *[a.ctypes.data for a in out]), self._h)
        return tuple(out)

    def read_verlet(self):
        n = self.info().n_particles
        out = [np.empty(n, np.float32) for _ in range(4)]
        _check(lib().pm_read_verlet(self._h, *[a.ctypes.data for a in out]), self._h)
        return tuple(out)

    def read_lattice(self, which):
        n = self.n
        if which == LATTICE_MASS:
            a = np.empty((n, n), np.int64)
        elif which == LATTICE_POTENTIAL:
            a = np.empty((n, n), np.float32)
        elif which == LATTICE_FORCE:
            a = np.empty((n, n, 2), np.float32)
        elif which == LATTICE_TRANSFER:
            a = np.empty((n // 2 + 1, n), np.float32)
        else:
            raise ValueError(which)
        _check(lib().pm_read_lattice(self._h, which, a.ctypes.data, a.nbytes), self._h)
        return a

    # stepping
    def step(self, k=1):
        _check(lib().pm_step(self._h, k), self._h)

    def stage(self, s):
        _check(lib().pm_stage(self._h, s), self._h)

    def sync(self):
        _check(lib().pm_sync(self._h), self._h)

    def frame(self, blur_r=3, reset=False):
        """Density frame of the current mass lattice, float32 [N, N] (index [y, x])."""
        n = self.cfg.n_lattice
        out = np.empty((n, n), np.float32)
        _check(lib().pm_frame(self._h, int(blur_r), 1 if reset else 0, out.ctypes.data), self._h)
        return out

    def diagnostics(self):
        out = (ctypes.c_double * 8)()
        _check(lib().pm_diagnostics(self._h, out), self._h)
        keys = ["mass", "px", "py", "kinetic", "potential", "max_speed", "nonfinite", "in_bounds"]
        return dict(zip(keys, list(out)))

    # timing
    def profile(self, on=True):
        _check(lib().pm_profile(self._h, 1 if on else 0), self._h)

    def profile_reset(self):
        _check(lib().pm_profile_reset(self._h), self._h)

    def profile_read(self):
        res = {}
        for k, name in enumerate(KERNELS):
            ms = ctypes.c_double()
            cnt = ctypes.c_int64()
            _check(lib().pm_profile_read(self._h, k, ctypes.byref(ms), ctypes.byref(cnt)), self._h)
            res[name] = (ms.value, cnt.value)
        return res


def restore(data, device=0):
    """A Simulation restored from checkpoint bytes."""
    buf = np.frombuffer(data, np.uint8)
    h = _H()
    _check(lib().pm_checkpoint_restore(buf.ctypes.data, buf.size, int(device), ctypes.byref(h)))
    return Simulation._adopt(h)


def open_checkpoint(path, device=0):
    """A Simulation restored from a checkpoint file."""
    h = _H()
    _check(lib().pm_checkpoint_open(os.fsencode(path), int(device), ctypes.byref(h)))
    return Simulation._adopt(h)


# ----- serving engine (include/pm_serve.h) -----------------------------------------

SERVE_PLAIN, SERVE_BATCHED = 0, 1
REQ_QUEUED, REQ_ACTIVE, REQ_DONE, REQ_FAILED, REQ_UNKNOWN = range(5)


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_int64),
        ("max_active", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("round_steps", ctypes.c_int32),
        ("use_graphs", ctypes.c_int32),
        ("batch_max_lattice", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("batch_max_particles", ctypes.c_int32),
        ("pinned_staging_mib", ctypes.c_int32),
        ("persistent", ctypes.c_int32),
        ("batch_tiled", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 4),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in ("submitted", "completed", "failed", "active", "queued")] + [
        (k, ctypes.c_double) for k in ("particle_steps", "steps", "first_submit_s", "last_done_s",
                                       "step_latency_p50_ms", "step_latency_p99_ms",
                                       "request_latency_p50_ms", "request_latency_p99_ms")] + [
        (k, ctypes.c_int64) for k in ("bytes_in_use", "peak_bytes", "rounds", "batched_launch_rounds",
                                      "graph_launches", "graph_captures", "pool_hits")] + [
        (k, ctypes.c_double) for k in ("admit_s", "launch_s", "readback_s", "wait_s", "create_s", "upload_s")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _declare_server(L):
    L.pm_server_config_default.argtypes = [ctypes.POINTER(ServerConfig)]
    L.pm_plan_admission.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.c_int32, ctypes.c_int64, ctypes.c_int32]
    L.pm_plan_admission.restype = ctypes.c_int32
    L.pm_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(_H)]
    L.pm_server_submit.argtypes = [_H, ctypes.POINTER(Config), _f32p, _f32p, _f32p, _f32p, _f32p, ctypes.c_int64,
                                   ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)]
    L.pm_server_status.argtypes = [_H, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32)]
    L.pm_server_wait.argtypes = [_H, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 ctypes.c_void_p, ctypes.c_double]
    L.pm_server_drain.argtypes = [_H]
    L.pm_server_get_stats.argtypes = [_H, ctypes.POINTER(ServerStats)]
    L.pm_server_destroy.argtypes = [_H]
    L.pm_server_last_error.argtypes = [_H]
    L.pm_server_last_error.restype = ctypes.c_char_p


def default_server_config(**kw):
    c = ServerConfig()
    lib().pm_server_config_default(ctypes.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


def plan_admission(need_bytes, free_bytes, max_new):
    """Number of queue-head requests admitted (strict FIFO), pure host logic."""
    a = (ctypes.c_int64 * max(1, len(need_bytes)))(*need_bytes)
    return lib().pm_plan_admission(a, len(need_bytes), int(free_bytes), int(max_new))


class Server:
    """Many independent simulations served concurrently on one device."""

    def __init__(self, cfg=None, **kw):
        self.cfg = cfg if cfg is not None else default_server_config(**kw)
        h = _H()
        st = lib().pm_server_create(ctypes.byref(self.cfg), ctypes.byref(h))
        if st != PM_OK:
            raise PMError(st, lib().pm_last_error(None).decode(errors="replace") or "pm_server_create failed")
        self._h = h
        self._n = {}

    def _check(self, st):
        if st != PM_OK:
            raise PMError(st, lib().pm_server_last_error(self._h).decode(errors="replace"))

    def submit(self, cfg, x, y, vx, vy, m, steps):
        arrs = [_f32(a) for a in (x, y, vx, vy, m)]
        n = arrs[0].size
        rid = ctypes.c_int64()
        self._check(lib().pm_server_submit(self._h, ctypes.byref(cfg), *arrs, n, int(steps), ctypes.byref(rid)))
        self._n[rid.value] = n
        return rid.value

    def status(self, rid):
        st = ctypes.c_int32()
        self._check(lib().pm_server_status(self._h, rid, ctypes.byref(st)))
        return st.value

    def wait(self, rid, timeout_s=0.0, copy=True):
        """Final (x, y, vx, vy) of a request, in load order (None with copy=False)."""
        n = self._n[rid]
        out = [np.empty(n, np.float32) for _ in range(4)] if copy else None
        ptrs = [o.ctypes.data for o in out] if copy else [None] * 4
        self._check(lib().pm_server_wait(self._h, rid, *ptrs, float(timeout_s)))
        del self._n[rid]
        return tuple(out) if copy else None

    def drain(self):
        self._check(lib().pm_server_drain(self._h))

    def stats(self):
        s = ServerStats()
        self._check(lib().pm_server_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            lib().pm_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
