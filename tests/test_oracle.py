"""CPU tests that pin the oracle (oracle/pm_oracle.c) before it is trusted as
the checker of the HIP path.  Every check is one a plausible slip (dropped
term, wrong sign, transposed index) would fail: brute force, closed forms,
invariants and numpy fixtures from tests/golden/make_golden.py."""
import os

import numpy as np
import pytest

from oracle import oracle as orc


def _load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


# ---------------------------------------------------------------- FFT --
@pytest.mark.parametrize("n", [8, 16, 32])
def test_fft1d_matches_brute_force_dft(n):
    rng = np.random.default_rng(n)
    z = rng.normal(size=n) + 1j * rng.normal(size=n)
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    assert np.allclose(orc.fft1d(z, -1), W @ z, atol=1e-12)
    assert np.allclose(orc.fft1d(z, +1), np.conj(W) @ z, atol=1e-12)


@pytest.mark.parametrize("n", [8, 16, 32])
def test_fft_matches_numpy_fixture(golden_dir, n):
    g = _load(golden_dir, f"fft_{n}.npz")
    assert np.allclose(orc.fft1d(g["z"], -1), g["fwd"], atol=1e-12)
    assert np.allclose(orc.fft1d(g["z"], +1), g["inv"], atol=1e-12)
    assert np.allclose(orc.fft2d(g["z2"], -1), g["fwd2"], atol=1e-11)


def test_fft2d_invariants():
    n = 64
    rng = np.random.default_rng(3)
    z = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    Z = orc.fft2d(z, -1)
    assert np.isclose(np.sum(np.abs(Z) ** 2), n * n * np.sum(np.abs(z) ** 2), rtol=1e-12)   # Parseval
    assert np.allclose(orc.fft2d(Z, +1), n * n * z, atol=1e-9)                              # round trip
    d = np.zeros((n, n)); d[0, 0] = 1.0
    assert np.allclose(orc.fft2d(d, -1), 1.0)                                             # delta -> constant
    s = np.roll(z, (3, 5), axis=(0, 1))                                                     # shift -> phase
    k = np.arange(n)
    ph = np.exp(-2j * np.pi * (5 * k[None, :] + 3 * k[:, None]) / n)
    assert np.allclose(orc.fft2d(s, -1), Z * ph, atol=1e-9)


# ------------------------------------------------------------- spectra --
def test_green_closed_form():
    n = 16
    G = orc.green(n, 3.0)
    assert G[0, 0] == 3.0
    assert G[0, 1] == 1.0 and G[1, 0] == 1.0
    assert np.isclose(G[3, 4], 0.2)                          # r = 5
    assert np.isclose(G[n - 3, n - 4], 0.2)                  # wrapped offsets (-3, -4)
    assert np.isclose(G[0, n // 2], 1.0 / (n // 2))          # i = N/2 is +N/2, not -N/2
    assert orc.green(n, 1.0)[0, 0] == 1.0


@pytest.mark.parametrize("n", [16, 32])
@pytest.mark.parametrize("mode", [0, 1])
def test_transfer_matches_numpy_fixture(golden_dir, n, mode):
    g = _load(golden_dir, f"transfer_{n}_m{mode}.npz")
    for key, g0 in (("T_g3", 3.0), ("T_g1", 1.0)):
        re, im = orc.transfer(n, mode, g0, 1e-3)
        assert np.allclose(re + 1j * im, g[key], atol=1e-10)


@pytest.mark.parametrize("mode", [0, 1])
def test_transfer_closed_form(mode):
    """NGP: T = Ghat (-1)^(kx+ky)/(1+eps).  CIC: T = Ghat s c/(c^2+eps),
    c = cos(pi kx_s/N) cos(pi ky_s/N).  Ghat is real because G is real-even."""
    n, eps = 32, 1e-3
    re, im = orc.transfer(n, mode, 1.0, eps)
    Gh = np.fft.fft2(orc.green(n, 1.0))
    assert np.abs(Gh.imag).max() < 1e-9
    k = np.arange(n)
    s = (-1.0) ** (k[None, :] + k[:, None])
    if mode == 0:
        c = np.ones((n, n))
    else:
        ks = np.where(k <= n // 2, k, k - n)
        c = np.cos(np.pi * ks[None, :] / n) * np.cos(np.pi * ks[:, None] / n)
    T = Gh.real * s * c / (c * c + eps)
    assert np.abs(im).max() < 1e-9
    assert np.allclose(re, T, atol=1e-9)


# ------------------------------------------------------------- deposit --
@pytest.mark.parametrize("mode", [0, 1])
def test_deposit_matches_brute_force_fixture(golden_dir, mode):
    g = _load(golden_dir, f"deposit_16_m{mode}.npz")
    rho = orc.deposit(16, mode, int(g["shift"]), g["x"], g["y"], g["m"])
    assert np.array_equal(rho, g["rho"])


def test_deposit_invariants():
    n, shift = 64, 30
    rng = np.random.default_rng(7)
    npart = 5000
    x = rng.uniform(0, n, npart).astype(np.float32)
    y = rng.uniform(0, n, npart).astype(np.float32)
    m = rng.lognormal(0, 0.5, npart).astype(np.float32)
    inb = (x >= 1) & (x < n - 1) & (y >= 1) & (y < n - 1)
    # NGP: exact sum of quantised masses, only in-bounds particles
    r0 = orc.deposit(n, 0, shift, x, y, m)
    q = np.rint(m[inb].astype(np.float64) * 2.0 ** shift).astype(np.int64)
    assert r0.sum() == q.sum()
    # CIC: partition of unity (within one rounding per contribution)
    r1 = orc.deposit(n, 1, shift, x, y, m)
    assert abs(int(r1.sum()) - int(q.sum())) <= 4 * inb.sum()
    # permutation invariance: bit-identical
    p = rng.permutation(npart)
    assert np.array_equal(orc.deposit(n, 1, shift, x[p], y[p], m[p]), r1)
    # bounds: x = N-1 exactly is out, x just below is in, x = 1 in, 0.999 out
    xs = np.float32([n - 1.0, np.nextafter(np.float32(n - 1), np.float32(0)), 1.0, 0.999])
    ys = np.float32([5.5] * 4)
    r = orc.deposit(n, 0, 0, xs, ys, np.ones(4, np.float32))
    assert r[5, n - 2] == 1 and r[5, 1] == 1 and r.sum() == 2
    # single CIC particle: weights (1-fx)(1-fy), fx(1-fy), (1-fx)fy, fx fy
    r = orc.deposit(n, 1, 8, np.float32([10.25]), np.float32([20.5]), np.float32([1.0]))
    assert r[20, 10] == 96 and r[20, 11] == 32 and r[21, 10] == 96 and r[21, 11] == 32


# ----------------------------------------------------------- potential --
@pytest.mark.parametrize("n,mode", [(16, 1), (32, 0), (32, 1)])
def test_potential_matches_numpy_fixture(golden_dir, n, mode):
    g = _load(golden_dir, f"potential_{n}_m{mode}.npz")
    T = orc.transfer(n, mode, 1.0, 1e-3)
    phi = orc.potential_d(n, T, g["rho"])
    assert np.allclose(phi, g["phi"], rtol=1e-10, atol=1e-12)


def test_potential_direct_convolution_ngp():
    """NGP: Phi = (rho (*) G)/(1+eps) as a direct O(N^4) periodic sum; the
    (-1)^k of the kernel spectrum and the N/2 output shift cancel."""
    n, eps = 8, 1e-3
    rng = np.random.default_rng(11)
    rho = rng.uniform(0, 1, (n, n))
    G = orc.green(n, 1.0)
    direct = np.zeros((n, n))
    for j in range(n):
        for i in range(n):
            for b in range(n):
                for a in range(n):
                    direct[j, i] += rho[b, a] * G[(j - b) % n, (i - a) % n]
    phi = orc.potential_d(n, orc.transfer(n, 0, 1.0, eps), rho)
    assert np.allclose(phi, direct / (1 + eps), rtol=1e-10)


def test_potential_single_particle_centred():
    """A unit mass in cell (j,i)=(20,37): the potential is mirror-symmetric
    about that cell (a shift slip of 1 or N/2 breaks it), and with the
    reference's origin value 3.0 the NGP peak sits on the cell itself."""
    n = 64
    rho = np.zeros((n, n)); rho[20, 37] = 1.0
    for mode in (0, 1):
        phi = orc.potential_d(n, orc.transfer(n, mode, 3.0, 1e-3), rho)
        p = np.roll(phi, (-20, -37), axis=(0, 1))
        assert np.allclose(p, p[::-1, :][np.r_[-1, :n - 1], :], atol=1e-12)   # j -> -j
        assert np.allclose(p, p[:, ::-1][:, np.r_[-1, :n - 1]], atol=1e-12)   # i -> -i
        assert np.allclose(p, p.T, atol=1e-12)
        if mode == 0:
            assert np.unravel_index(np.argmax(phi), phi.shape) == (20, 37)


# ------------------------------------------------------------ gradient --
def test_gradient_cubic_exact_and_edge_rule():
    n = 32
    j, i = np.meshgrid(np.arange(n, dtype=float), np.arange(n, dtype=float), indexing="ij")
    phi = 0.01 * i ** 3 - 0.2 * i * j + 0.03 * j ** 3 + 0.5 * j ** 2
    fx, fy = orc.gradient64(n, 2.0, phi)
    ex = 2.0 * (0.03 * i ** 2 - 0.2 * j)
    ey = 2.0 * (-0.2 * i + 0.09 * j ** 2 + 1.0 * j)
    inner = np.s_[2:n - 2, 2:n - 2]
    assert np.allclose(fx[inner], ex[inner], atol=1e-10)
    assert np.allclose(fy[inner], ey[inner], atol=1e-10)
    # edge: at i = 0 neighbours i-1, i-2 are replaced by the centre value
    c = phi[5, 0]
    want = (-phi[5, 2] + 8 * phi[5, 1] - 8 * c + c) / 12 * 2.0
    assert np.isclose(fx[5, 0], want)
    c = phi[n - 2, 7]
    want = (-c + 8 * phi[n - 1, 7] - 8 * phi[n - 3, 7] + phi[n - 4, 7]) / 12 * 2.0
    assert np.isclose(fy[n - 2, 7], want)


def test_gradient32_matches_gradient64():
    n = 32
    rng = np.random.default_rng(5)
    phi = rng.normal(size=(n, n)).astype(np.float32)
    f32 = orc.gradient32(n, 1.5, phi)
    fx, fy = orc.gradient64(n, 1.5, phi.astype(np.float64))
    assert np.allclose(f32[..., 0], fx, atol=1e-5) and np.allclose(f32[..., 1], fy, atol=1e-5)


# ---------------------------------------------------------------- push --
def test_push_rest_start_and_linear_field():
    n, dt = 64, 0.1
    j, i = np.meshgrid(np.arange(n, dtype=np.float32), np.arange(n, dtype=np.float32), indexing="ij")
    fxy = np.stack([0.5 + 0.25 * i, -0.125 * j], axis=-1).astype(np.float32)
    x = np.float32([10.5, 20.25, 30.0])
    y = np.float32([11.5, 40.75, 33.0])
    z = np.zeros(3, np.float32)
    for mode in (0, 1):
        nx, ny, vx, vy = orc.push32(n, mode, dt, 0, x, y, z, z, fxy)
        xs = np.floor(x) if mode == 0 else x          # NGP samples the cell corner, CIC is exact on linear fields
        ys = np.floor(y) if mode == 0 else y
        ax = 0.5 + 0.25 * xs
        ay = -0.125 * ys
        assert np.allclose(vx, ax * dt, rtol=1e-6) and np.allclose(vy, ay * dt, rtol=1e-6)
        assert np.allclose(nx, x + ax * dt * dt, rtol=1e-6) and np.allclose(ny, y + ay * dt * dt, rtol=1e-6)


def test_push_wrap_and_bounds():
    n, dt = 64, 1.0
    fxy = np.zeros((n, n, 2), np.float32)
    x = np.float32([2.5, 61.5, 0.5, 30.0])
    y = np.float32([30.0, 30.0, 30.0, 30.0])
    vx = np.float32([-1.0, 0.75, 5.0, 0.0])
    vy = np.zeros(4, np.float32)
    nx, ny, nvx, nvy = orc.push32(n, 1, dt, 0, x, y, vx, vy, fxy)
    assert nx[0] == np.float32(1.5 + 60) and nvx[0] == -1.0      # wrapped, velocity kept
    assert nx[1] == np.float32(62.25 - 60) and nvx[1] == 0.75
    assert nx[2] == x[2] and nvx[2] == vx[2]                      # out of bounds: untouched
    nx, _, nvx, _ = orc.push32(n, 1, dt, 1, x, y, vx, vy, fxy)
    assert nvx[0] == np.float32(-1.0 + 60.0)                      # reference quirk: implied velocity jumps


# --------------------------------------------------------- trajectory --
@pytest.mark.parametrize("mode", [0, 1])
def test_step64_matches_numpy_trajectory(golden_dir, mode):
    g = _load(golden_dir, f"traj64_32_m{mode}.npz")
    s = orc.Step64(32, mode, float(g["dt"]), float(g["g"]), 1.0, 1e-3, 0,
                   g["x0"], g["y0"], g["vx0"], g["vy0"], g["m"])
    s.step(5)
    for k in ("x", "y", "vx", "vy"):
        assert np.allclose(getattr(s, k), g[k], rtol=1e-9, atol=1e-9)
    assert np.allclose(s.phi.reshape(32, 32), g["phi_last"], rtol=1e-9, atol=1e-12)


def test_two_body_symmetry_and_momentum():
    """Two equal masses placed symmetrically about the lattice centre: equal
    and opposite accelerations, total momentum stays zero (CIC deposit + CIC
    gather + antisymmetric stencil)."""
    n = 64
    # isolated particle on a cell corner: zero self-force (symmetric potential, stencil antisymmetric)
    for mode in (0, 1):
        s = orc.Step64(n, mode, 0.05, 1.0, 1.0, 1e-3, 0, np.array([21.0]), np.array([40.0]),
                       np.zeros(1), np.zeros(1), np.ones(1))
        s.step(1)
        assert abs(s.vx[0]) < 1e-12 and abs(s.vy[0]) < 1e-12
    x = np.array([20.0, 44.0]); y = np.array([30.0, 34.0])
    z = np.zeros(2)
    s = orc.Step64(n, 1, 0.05, 1.0, 1.0, 1e-3, 0, x, y, z, z, np.ones(2))
    s.step(1)
    assert np.allclose(s.vx[0], -s.vx[1], rtol=1e-9) and np.allclose(s.vy[0], -s.vy[1], rtol=1e-9)
    assert s.vx[0] > 0 and s.vy[0] > 0                           # attractive
    s.step(10)
    assert abs(s.vx.sum()) < 1e-9 * abs(s.vx).max() + 1e-12


def test_mixed_step_tracks_fp64():
    """FP32-state step with int64 deposit vs the all-FP64 step on a C1-like case."""
    n, npart = 64, 4000
    rng = np.random.default_rng(9)
    x = rng.uniform(16, 48, npart).astype(np.float32)
    y = rng.uniform(16, 48, npart).astype(np.float32)
    z = np.zeros(npart, np.float32)
    m = np.ones(npart, np.float32)
    a = orc.MixedStepper(n, 1, 0.01, 1.0, 1.0, 1e-3, 0, 40, x, y, z, z, m)
    b = orc.Step64(n, 1, 0.01, 1.0, 1.0, 1e-3, 0, x, y, z, z, m)
    a.step(5); b.step(5)
    assert np.abs(a.x - b.x).max() < 1e-3 and np.abs(a.y - b.y).max() < 1e-3


@pytest.mark.parametrize("mode", [0, 1])
def test_cpu32_step_matches_oracle(mode):
    """The FP32 CPU step (the CPU baseline): its potential is the oracle's FP64
    potential of the same int64 deposit within 1e-5 relative L2 (FP32 FFT
    rounding), and its trajectory tracks the FP64-field step."""
    n, npart, shift = 64, 4000, 40
    rng = np.random.default_rng(11 + mode)
    x = rng.uniform(16, 48, npart).astype(np.float32)
    y = rng.uniform(16, 48, npart).astype(np.float32)
    z = np.zeros(npart, np.float32)
    m = rng.lognormal(0.0, 0.5, npart).astype(np.float32)
    c = orc.Cpu32Stepper(n, mode, 0.01, 1.0, 1.0, 1e-3, 0, shift, x, y, z, z, m)
    c.step(1)
    phi64 = orc.potential(n, orc.transfer(n, mode, 1.0, 1e-3), shift, orc.deposit(n, mode, shift, x, y, m))
    phi32 = c.potential().astype(np.float64)
    assert np.linalg.norm(phi32 - phi64) / np.linalg.norm(phi64) < 1e-5
    a = orc.MixedStepper(n, mode, 0.01, 1.0, 1.0, 1e-3, 0, shift, x, y, z, z, m)
    a.step(1)
    c.step(4)
    a.step(4)
    assert np.abs(c.x - a.x).max() < 1e-3 and np.abs(c.y - a.y).max() < 1e-3
