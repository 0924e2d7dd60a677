"""GPU tests of the density-frame stage (pm_frame), which replaces the
reference's render pipeline (CosmosCuda.cuh:613-828, 1249-1256).

Pinned by closed forms, not by retyping the kernel:
  * with the reference's BLUR_R = 3 the filter keeps only the centre weight
    1/(0+1) (every other offset has r >= 1), so the frame is
    1.8 * ((mass - min) / (max - min))^0.25 of the per-cell mass; the heaviest
    cell maps to exactly 1.8 and empty cells to 0;
  * with blur 5 a unit mass spreads with weights 1 (centre), 1/2 (edge
    neighbours), 1/(1+sqrt 2) (diagonals) and 0 beyond;
  * the extremes run over frames until reset.
Tolerance: 2 ulp-level relative error (powf vs numpy power in float32).
"""
import numpy as np
import pytest

from cosmossimulationwithcuda_amd import pm

pytestmark = pytest.mark.gpu


def ngp_sim(n, x, y, m):
    cfg = pm.default_config(n_lattice=n, n_particles=len(x), mode=pm.NGP, dt=0.01, green_origin=1.0,
                            deconv_eps=1e-3)
    s = pm.Simulation(cfg)
    z = np.zeros(len(x), np.float32)
    s.load(np.float32(x), np.float32(y), z, z, np.float32(m))
    return s


def test_frame_identity_blur_is_scaled_mass():
    n = 256
    rng = np.random.default_rng(5)
    k = 3000
    cx = rng.integers(2, n - 2, k)
    cy = rng.integers(2, n - 2, k)
    m = rng.lognormal(0, 0.5, k).astype(np.float32)
    with ngp_sim(n, cx + 0.5, cy + 0.5, m) as s:
        frame = s.frame(blur_r=3, reset=True)
        rho = s.read_lattice(pm.LATTICE_MASS)
        shift = s.info().mass_shift
    mass = rho.astype(np.float32) * np.float32(2.0 ** -shift)
    lo, hi = mass.min(), mass.max()
    want = np.float32(1.8) * np.power((mass - lo) / (hi - lo), np.float32(0.25))
    assert frame.shape == (n, n)
    assert frame.max() == np.float32(1.8) and frame[mass == lo].max() == 0.0
    np.testing.assert_allclose(frame, want, rtol=3e-7, atol=1e-7)


def test_frame_blur5_kernel_weights():
    n = 128
    with ngp_sim(n, [60.5], [70.5], [1.0]) as s:
        frame = s.frame(blur_r=5, reset=True)
    # max = 1 (centre), min = 0: frame = 1.8 * w^0.25
    w = {0: 1.0, 1: 0.5, 2: 1.0 / (1.0 + np.sqrt(2.0))}
    for dy in range(-3, 4):
        for dx in range(-3, 4):
            d2 = dx * dx + dy * dy
            expect = 1.8 * w[d2] ** 0.25 if d2 in w else 0.0
            assert abs(frame[70 + dy, 60 + dx] - expect) <= 2e-6, (dx, dy)
    assert np.count_nonzero(frame) == 9


def test_frame_running_extremes():
    n = 128
    x, y, m = [30.5, 90.5], [40.5, 80.5], [4.0, 1.0]
    with ngp_sim(n, x, y, m) as s:
        f1 = s.frame(reset=True)
        assert f1[40, 30] == np.float32(1.8)
        z = np.zeros(2, np.float32)
        s.load(np.float32(x), np.float32(y), z, z, np.float32([1.0, 1.0]))   # lighter frame
        f2 = s.frame()                           # extremes from frame 1 still hold
        f3 = s.frame(reset=True)                 # restarted
    # frame 2: both cells 1.0 on a running range [0, 4] (shift may differ: compare values)
    assert abs(f2[40, 30] - 1.8 * 0.25 ** 0.25) <= 2e-6 and abs(f2[80, 90] - f2[40, 30]) <= 2e-6
    assert f3[40, 30] == np.float32(1.8) and f3[80, 90] == np.float32(1.8)
