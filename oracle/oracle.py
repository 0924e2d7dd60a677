"""ctypes front end of the CPU oracle (pm_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product library never loads
it.  Each wrapper names the reference lines its C function restates
(CosmosCuda.cuh abbreviated "cuh").
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpmoracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_c = ctypes


def build():
    """Compile pm_oracle.c with the Makefile beside it."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_fft1d.argtypes = [_c.c_int, _f64p, _f64p, _c.c_int]
        L.orc_fft2d.argtypes = [_c.c_int, _f64p, _f64p, _c.c_int]
        L.orc_green.argtypes = [_c.c_int, _c.c_double, _f64p]
        L.orc_mass_kernel.argtypes = [_c.c_int, _c.c_int, _f64p]
        L.orc_transfer.argtypes = [_c.c_int, _c.c_int, _c.c_double, _c.c_double, _f64p, _f64p]
        L.orc_deposit.argtypes = [_c.c_int, _c.c_int, _c.c_int, _c.c_int64, _f32p, _f32p, _f32p, _i64p]
        L.orc_deposit64.argtypes = [_c.c_int, _c.c_int, _c.c_int64, _f64p, _f64p, _f64p, _f64p]
        L.orc_potential.argtypes = [_c.c_int, _f64p, _f64p, _c.c_int, _i64p, _f64p]
        L.orc_potential_d.argtypes = [_c.c_int, _f64p, _f64p, _f64p, _f64p]
        L.orc_gradient64.argtypes = [_c.c_int, _c.c_double, _f64p, _f64p, _f64p]
        L.orc_gradient32.argtypes = [_c.c_int, _c.c_float, _f32p, _f32p]
        L.orc_push32.argtypes = [_c.c_int, _c.c_int, _c.c_float, _c.c_int, _c.c_int64,
                                 _f32p, _f32p, _f32p, _f32p, _f32p]
        L.orc_push64.argtypes = [_c.c_int, _c.c_int, _c.c_double, _c.c_int, _c.c_int64,
                                 _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.orc_step_mixed.argtypes = [_c.c_int, _c.c_int, _c.c_float, _c.c_float, _c.c_int, _c.c_int,
                                     _f64p, _f64p, _c.c_int64, _f32p, _f32p, _f32p, _f32p, _f32p,
                                     _i64p, _f64p, _f64p, _f64p, _f32p]
        L.orc_step64.argtypes = [_c.c_int, _c.c_int, _c.c_double, _c.c_double, _c.c_int,
                                 _f64p, _f64p, _c.c_int64, _f64p, _f64p, _f64p, _f64p, _f64p,
                                 _f64p, _f64p, _f64p, _f64p]
        L.cpu32_create.argtypes = [_c.c_int, _c.c_int, _c.c_float, _c.c_float, _c.c_int, _c.c_int, _f64p, _f64p]
        L.cpu32_create.restype = _c.c_void_p
        L.cpu32_destroy.argtypes = [_c.c_void_p]
        L.cpu32_step.argtypes = [_c.c_void_p, _c.c_int64, _f32p, _f32p, _f32p, _f32p, _f32p, _c.c_int]
        L.cpu32_potential.argtypes = [_c.c_void_p]
        L.cpu32_potential.restype = _c.POINTER(_c.c_float)
        L.orc_num_threads.restype = _c.c_int
        L.orc_set_threads.argtypes = [_c.c_int]
        _lib = L
    return _lib


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def set_threads(t):
    lib().orc_set_threads(int(t))


def num_threads():
    return lib().orc_num_threads()


def fft1d(z, sign=-1):
    """cufftExecC2C restated (cuh:1266-1294): unnormalised, sign -1 forward."""
    re, im = _c64(np.real(z)).copy(), _c64(np.imag(z)).copy()
    assert lib().orc_fft1d(len(re), re, im, sign) == 0
    return re + 1j * im


def fft2d(z, sign=-1):
    """2-D transform, index [j, i] (y rows, x fastest)."""
    n = z.shape[0]
    re, im = _c64(np.real(z)).copy(), _c64(np.imag(z)).copy()
    assert lib().orc_fft2d(n, re, im, sign) == 0
    return re + 1j * im


def green(n, g0):
    """Periodic 1/r lattice, cuh:89-119.  Returns [j, i]."""
    G = np.empty((n, n), np.float64)
    lib().orc_green(n, g0, G)
    return G


def mass_kernel(n, mode):
    """Spatial assignment kernel, cuh:374-415."""
    D = np.empty((n, n), np.float64)
    lib().orc_mass_kernel(n, mode, D)
    return D


def transfer(n, mode, g0, eps):
    """Composite spectral factor Ghat*conj(Dhat)/(|Dhat|^2+eps) (cuh:120-173,
    416-445, 1018-1027) as complex [j, i]."""
    re = np.empty((n, n), np.float64)
    im = np.empty((n, n), np.float64)
    lib().orc_transfer(n, mode, g0, eps, re, im)
    return re, im


def deposit(n, mode, shift, x, y, m):
    """int64 fixed-point deposit, cuh:446-595.  Returns rho[j, i] int64."""
    x, y, m = _c32(x), _c32(y), _c32(m)
    rho = np.empty((n, n), np.int64)
    lib().orc_deposit(n, mode, shift, len(x), x, y, m, rho)
    return rho


def deposit64(n, mode, x, y, m):
    x, y, m = _c64(x), _c64(y), _c64(m)
    rho = np.empty((n, n), np.float64)
    lib().orc_deposit64(n, mode, len(x), x, y, m, rho)
    return rho


def potential(n, T, shift, rho):
    """FFT -> product -> IFFT -> shift/scale, cuh:1245-1297, FP64."""
    phi = np.empty((n, n), np.float64)
    lib().orc_potential(n, _c64(T[0]), _c64(T[1]), shift, np.ascontiguousarray(rho, np.int64), phi)
    return phi


def potential_d(n, T, rho_d):
    phi = np.empty((n, n), np.float64)
    lib().orc_potential_d(n, _c64(T[0]), _c64(T[1]), _c64(rho_d), phi)
    return phi


def gradient64(n, g, phi):
    """4th-order stencil with centre clamp, cuh:174-229.  Returns (Fx, Fy)."""
    fx = np.empty((n, n), np.float64)
    fy = np.empty((n, n), np.float64)
    lib().orc_gradient64(n, g, _c64(phi), fx, fy)
    return fx, fy


def gradient32(n, g, phi32):
    """FP32 mirror of the product stencil.  Returns interleaved [j, i, 2]."""
    out = np.empty((n, n, 2), np.float32)
    lib().orc_gradient32(n, g, _c32(phi32), out)
    return out


def push32(n, mode, dt, wrap_mode, x, y, vx, vy, fxy):
    """FP32 gather + kick-drift mirror, cuh:231-312.  Returns new arrays."""
    x, y, vx, vy = (_c32(a).copy() for a in (x, y, vx, vy))
    lib().orc_push32(n, mode, dt, wrap_mode, len(x), x, y, vx, vy, _c32(fxy))
    return x, y, vx, vy


def push64(n, mode, dt, wrap_mode, x, y, vx, vy, fx, fy):
    x, y, vx, vy = (_c64(a).copy() for a in (x, y, vx, vy))
    lib().orc_push64(n, mode, dt, wrap_mode, len(x), x, y, vx, vy, _c64(fx), _c64(fy))
    return x, y, vx, vy


class MixedStepper:
    """FP32 particle state, int64 deposit, FP64 field: the CPU implementation
    of the per-step algorithm (cuh:1336-1341 order)."""

    def __init__(self, n, mode, dt, g, g0, eps, wrap_mode, shift, x, y, vx, vy, m):
        self.n, self.mode, self.dt, self.g = n, mode, dt, g
        self.wrap_mode, self.shift = wrap_mode, shift
        self.T = tuple(_c64(a) for a in transfer(n, mode, g0, eps))
        self.x, self.y, self.vx, self.vy, self.m = (_c32(a).copy() for a in (x, y, vx, vy, m))
        nn = n * n
        self.rho = np.empty(nn, np.int64)
        self.phi = np.empty(nn, np.float64)
        self.fx = np.empty(nn, np.float64)
        self.fy = np.empty(nn, np.float64)
        self.fxy = np.empty(2 * nn, np.float32)

    def step(self, k=1):
        L = lib()
        for _ in range(k):
            L.orc_step_mixed(self.n, self.mode, self.dt, self.g, self.wrap_mode, self.shift,
                             self.T[0], self.T[1], len(self.x), self.x, self.y, self.vx, self.vy,
                             self.m, self.rho, self.phi, self.fx, self.fy, self.fxy)


class Step64:
    """All-FP64 trajectory oracle."""

    def __init__(self, n, mode, dt, g, g0, eps, wrap_mode, x, y, vx, vy, m):
        self.n, self.mode, self.dt, self.g, self.wrap_mode = n, mode, dt, g, wrap_mode
        self.T = tuple(_c64(a) for a in transfer(n, mode, g0, eps))
        self.x, self.y, self.vx, self.vy, self.m = (_c64(a).copy() for a in (x, y, vx, vy, m))
        nn = n * n
        self.rho = np.empty(nn, np.float64)
        self.phi = np.empty(nn, np.float64)
        self.fx = np.empty(nn, np.float64)
        self.fy = np.empty(nn, np.float64)

    def step(self, k=1):
        L = lib()
        for _ in range(k):
            L.orc_step64(self.n, self.mode, self.dt, self.g, self.wrap_mode, self.T[0], self.T[1],
                         len(self.x), self.x, self.y, self.vx, self.vy, self.m,
                         self.rho, self.phi, self.fx, self.fy)


class Cpu32Stepper:
    """The per-step algorithm in FP32 on the host cores (pm_cpu32.c): int64
    deposit into per-thread lattices allocated once, FP32 FFT convolution,
    the reference's FP32 stencil and push.  bench.py times it as the CPU
    baseline."""

    def __init__(self, n, mode, dt, g, g0, eps, wrap_mode, shift, x, y, vx, vy, m):
        self.n = n
        T = tuple(_c64(a) for a in transfer(n, mode, g0, eps))
        self.x, self.y, self.vx, self.vy, self.m = (_c32(a).copy() for a in (x, y, vx, vy, m))
        self._h = lib().cpu32_create(n, mode, float(dt), float(g), wrap_mode, shift, T[0], T[1])
        if not self._h:
            raise MemoryError("cpu32_create failed")

    def step(self, k=1):
        lib().cpu32_step(self._h, len(self.x), self.x, self.y, self.vx, self.vy, self.m, k)

    def potential(self):
        p = lib().cpu32_potential(self._h)
        return np.ctypeslib.as_array(p, shape=(self.n * self.n,)).reshape(self.n, self.n).copy()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().cpu32_destroy(self._h)
            self._h = None
