"""pm_diagnostics_noalloc: the diagnostics reductions over caller-owned device
arrays with the partial sums in the library's fixed device array (no device
allocation by the library).  Checked through the C ABI against the existing
entry point pm_diagnostics and against numpy FP64 sums, on two calls in a row,
the second with a smaller input.

Bars: max |v|, non-finite and in-bounds counts exact, and the mass sum where
it is exact in FP64 (equal masses); mass (lognormal masses), momentum and
kinetic energy within 1e-12 relative of pm_diagnostics (which
sums the same values in the handle's tile order) and of numpy; the potential
energy over the same lattice cells bit-exact against pm_diagnostics (same
kernel, same order); repeated calls bit-identical.
"""
import ctypes

import numpy as np
import pytest

from cosmossimulationwithcuda_amd import ics, pm

pytestmark = pytest.mark.gpu


class DeviceBuffers:
    """Device arrays from the HIP runtime libpm.so itself links (torch ships
    its own runtime; two runtimes in one process do not share a device)."""

    def __init__(self):
        pm.lib()
        self.hip = ctypes.CDLL("libamdhip64.so.7")      # the already-loaded runtime of libpm.so
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 1))) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1) == 0
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def call_noalloc(n, arrs, rho=None, phi=None, shift=0):
    d = DeviceBuffers()
    try:
        t = [d.put(a) for a in arrs]
        out = d.put(np.full(8, -1.0))
        tr = d.put(rho) if rho is not None else None
        tp = d.put(phi) if phi is not None else None
        st = pm.lib().pm_diagnostics_noalloc(0, None, n, *t, len(arrs[0]), tr, shift, tp, out)
        assert st == pm.PM_OK, pm.lib().pm_last_error(None)
        return d.get(out, 8, np.float64)
    finally:
        d.free()


def numpy_diag(n, arrs):
    x, y, vx, vy, m = (a.astype(np.float64) for a in arrs)
    fin = np.isfinite(x) & np.isfinite(y) & np.isfinite(vx) & np.isfinite(vy) & np.isfinite(m)
    x, y, vx, vy, m = x[fin], y[fin], vx[fin], vy[fin], m[fin]
    inb = (x >= 1) & (x < n - 1) & (y >= 1) & (y < n - 1)
    return np.array([m.sum(), (m * vx).sum(), (m * vy).sum(), 0.5 * (m * (vx * vx + vy * vy)).sum(), 0.0,
                     np.sqrt(vx * vx + vy * vy).max(), float((~fin).sum()), float(inb.sum())])


def close(a, b, scale):
    return abs(a - b) <= 1e-12 * scale


def test_noalloc_diagnostics_match_pm_diagnostics():
    # first call: 300k particles on 512^2 with the mass lattice and potential
    n, k = 512, 300_000
    x, y, vx, vy, m = ics.exponential_disk(k, n / 2, n / 2, n / 12, n / 3, seed=11, masses="lognormal")
    x[:5] = np.float32([0.5, n - 0.5, np.nan, 3.0, 700.0])      # out of bounds, non-finite
    arrs = (x, y, vx, vy, m)
    cfg = pm.default_config(n_lattice=n, n_particles=k, mode=pm.CIC, dt=0.01, gravity=1.0, green_origin=1.0,
                            deconv_eps=1e-3)
    with pm.Simulation(cfg) as s:
        s.load(*arrs)
        s.stage(pm.STAGE_DEPOSIT)
        s.stage(pm.STAGE_FIELD)
        want = s.diagnostics()
        rho = s.read_lattice(pm.LATTICE_MASS)
        phi = s.read_lattice(pm.LATTICE_POTENTIAL)
        shift = s.info().mass_shift
    print(f"first call: {k} particles, {n}^2 lattice", flush=True)
    got = call_noalloc(n, arrs, rho, phi, shift)
    ref = numpy_diag(n, arrs)
    pscale = float(np.nansum(np.abs(m.astype(np.float64)) * np.hypot(vx, vy)))
    assert close(got[0], want["mass"], want["mass"]) and got[5] == want["max_speed"]
    assert got[6] == want["nonfinite"] == 1 and got[7] == want["in_bounds"]
    assert close(got[1], want["px"], pscale) and close(got[2], want["py"], pscale)
    assert close(got[3], want["kinetic"], abs(want["kinetic"]))
    assert got[4] == want["potential"]                            # same cells, same kernel, same order
    assert close(got[0], ref[0], ref[0]) and got[5] == ref[5] and got[7] == ref[7]
    assert close(got[1], ref[1], pscale) and close(got[3], ref[3], ref[3])
    again = call_noalloc(n, arrs, rho, phi, shift)
    assert np.array_equal(again.view(np.uint64), got.view(np.uint64))

    # second call: a smaller input (50k particles, 256^2, no lattice); the
    # scratch still holds the first call's larger partial sums
    n2, k2 = 256, 50_000
    arrs2 = ics.uniform_square(k2, 32, 224, seed=5, mass=2.0)
    arrs2 = (arrs2[0], arrs2[1], np.full(k2, 0.25, np.float32), np.full(k2, -0.5, np.float32), arrs2[4])
    print(f"second call: {k2} particles, {n2}^2 lattice, no lattice terms", flush=True)
    got2 = call_noalloc(n2, arrs2)
    ref2 = numpy_diag(n2, arrs2)
    with pm.Simulation(pm.default_config(n_lattice=n2, n_particles=k2, mode=pm.CIC, dt=0.01)) as s:
        s.load(*arrs2)
        want2 = s.diagnostics()
    assert got2[0] == want2["mass"] == 2.0 * k2 and got2[4] == 0.0
    assert got2[6] == 0 and got2[7] == want2["in_bounds"] == k2 and got2[5] == want2["max_speed"]
    assert close(got2[1], ref2[1], abs(ref2[1])) and close(got2[2], ref2[2], abs(ref2[2]))
    assert close(got2[3], want2["kinetic"], want2["kinetic"])


def test_noalloc_rejects_bad_arguments():
    d = DeviceBuffers()
    o = d.put(np.zeros(8))
    L = pm.lib()
    assert L.pm_diagnostics_noalloc(0, None, 100, None, None, None, None, None, 0, None, 0, None, o) == pm.PM_EINVAL
    assert L.pm_diagnostics_noalloc(0, None, 256, None, None, None, None, None, 0, None, 0, None, None) == pm.PM_EINVAL
    assert L.pm_diagnostics_noalloc(0, None, 256, None, None, None, None, None, 10, None, 0, None, o) == pm.PM_EINVAL
    assert L.pm_diagnostics_noalloc(99, None, 256, None, None, None, None, None, 0, None, 0, None, o) == pm.PM_ENODEV
    d.free()
