"""Writes the golden fixtures under tests/golden/ (run: python tests/golden/make_golden.py).

The reference ships no tests, fixtures or golden vectors, and it cannot be
built here (CUDA + cuFFT + OpenCV, Windows MSBuild).  These fixtures are a
second, independent restatement of the reference's algorithm written with
numpy (numpy.fft for transforms, explicit per-particle loops for the deposit
and push), so they pin the C oracle through a different code path:

  fft_*.npz        numpy.fft FP64 forward/inverse vectors, N = 8, 16, 32
  transfer_*.npz   Ghat*conj(Dhat)/(|Dhat|^2+eps) built literally as the reference
                   does (cuh:89-173, 374-445, 1018-1027) with numpy.fft
  deposit_*.npz    brute-force per-particle int64 deposit (cuh:446-595 rule)
  potential_*.npz  FP64 potential for a fixed lattice (cuh:1245-1297)
  traj64_*.npz     5-step all-FP64 trajectory, N = 32 (cuh:1336-1341 order)
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def green(n, g0):
    i = np.arange(n)
    d = np.where(i <= n // 2, i, i - n).astype(np.float64)
    r = np.sqrt(d[:, None] ** 2 + d[None, :] ** 2)       # [j, i]
    out = np.full((n, n), g0)
    np.divide(1.0, r, out=out, where=r > 0.5)
    return out


def mass_kernel(n, mode):
    D = np.zeros((n, n))
    c = n // 2
    if mode == 0:
        D[c, c] = 1.0
    else:
        D[c:c + 2, c:c + 2] = 0.25
    return D


def transfer(n, mode, g0, eps):
    Gh = np.fft.fft2(green(n, g0))
    Dh = np.fft.fft2(mass_kernel(n, mode))
    if mode == 1:
        k = np.arange(n)
        ks = np.where(k <= n // 2, k, k - n).astype(np.float64)
        ph = -2.0 * np.pi * (ks[None, :] * -0.5 / n + ks[:, None] * -0.5 / n)
        Dh = Dh * np.exp(1j * ph)
    return Gh * np.conj(Dh) / (np.abs(Dh) ** 2 + eps)


def deposit_brute(n, mode, shift, x, y, m):
    rho = np.zeros((n, n), np.int64)
    scale = np.float32(2.0 ** shift)
    one = np.float32(1.0)
    for px, py, pm in zip(x.astype(np.float32), y.astype(np.float32), m.astype(np.float32)):
        if not (px >= 1 and px < n - 1 and py >= 1 and py < n - 1):
            continue
        xi, yi = int(px), int(py)
        if mode == 0:
            rho[yi, xi] += int(np.rint(pm * scale))
            continue
        fx = np.float32(px - np.float32(xi))
        fy = np.float32(py - np.float32(yi))
        ax, ay = np.float32(one - fx), np.float32(one - fy)
        for (dx, dy, w) in ((0, 0, ax * ay), (1, 0, fx * ay), (0, 1, ax * fy), (1, 1, fx * fy)):
            p = np.float32(np.float32(w) * pm)
            rho[yi + dy, xi + dx] += int(np.rint(np.float32(p * scale)))
    return rho


def potential(n, T, rho_d):
    # numpy's ifft2 already carries the 1/N^2 the reference applies after its unnormalised inverse
    L = np.fft.ifft2(np.fft.fft2(rho_d) * T).real
    return np.roll(L, (n // 2, n // 2), axis=(0, 1))        # Phi[(j+N/2)%N, (i+N/2)%N] = L[j, i]


def gradient(phi, g):
    n = phi.shape[0]
    fx = np.empty_like(phi)
    fy = np.empty_like(phi)
    for j in range(n):
        for i in range(n):
            p0 = phi[j, i]

            def at(jj, ii):
                return phi[jj, ii] if (0 <= ii < n and 0 <= jj < n) else p0
            fx[j, i] = (-at(j, i + 2) + 8 * at(j, i + 1) - 8 * at(j, i - 1) + at(j, i - 2)) / 12 * g
            fy[j, i] = (-at(j + 2, i) + 8 * at(j + 1, i) - 8 * at(j - 1, i) + at(j - 2, i)) / 12 * g
    return fx, fy


def step64(n, mode, dt, g, T, x, y, vx, vy, m):
    rho = np.zeros((n, n))
    for p in range(len(x)):
        px, py = x[p], y[p]
        if not (px >= 1 and px < n - 1 and py >= 1 and py < n - 1):
            continue
        xi, yi = int(px), int(py)
        ex, ey = px - xi, py - yi
        if mode == 0:
            rho[yi, xi] += m[p]
        else:
            rho[yi, xi] += (1 - ex) * (1 - ey) * m[p]
            rho[yi, xi + 1] += ex * (1 - ey) * m[p]
            rho[yi + 1, xi] += (1 - ex) * ey * m[p]
            rho[yi + 1, xi + 1] += ex * ey * m[p]
    phi = potential(n, T, rho)
    fx, fy = gradient(phi, g)
    for p in range(len(x)):
        px, py = x[p], y[p]
        if not (px >= 1 and px < n - 1 and py >= 1 and py < n - 1):
            continue
        xi, yi = int(px), int(py)
        ex, ey = px - xi, py - yi
        if mode == 0:
            ax, ay = fx[yi, xi], fy[yi, xi]
        else:
            w = ((0, 0, (1 - ex) * (1 - ey)), (1, 0, ex * (1 - ey)), (0, 1, (1 - ex) * ey), (1, 1, ex * ey))
            ax = sum(fx[yi + b, xi + a] * c for a, b, c in w)
            ay = sum(fy[yi + b, xi + a] * c for a, b, c in w)
        vx[p] += ax * dt
        vy[p] += ay * dt
        nx, ny = px + vx[p] * dt, py + vy[p] * dt
        span = n - 4.0
        if nx < 2:
            nx += span
        if ny < 2:
            ny += span
        if nx > n - 2:
            nx -= span
        if ny > n - 2:
            ny -= span
        x[p], y[p] = nx, ny
    return rho, phi, fx, fy


def main():
    rng = np.random.default_rng(20251203)
    for n in (8, 16, 32):
        z = rng.normal(size=n) + 1j * rng.normal(size=n)
        z2 = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        np.savez_compressed(os.path.join(HERE, f"fft_{n}.npz"), z=z, fwd=np.fft.fft(z),
                            inv=np.fft.ifft(z) * n, z2=z2, fwd2=np.fft.fft2(z2))
    for n in (16, 32):
        for mode in (0, 1):
            T = transfer(n, mode, 3.0, 1e-3)
            T1 = transfer(n, mode, 1.0, 1e-3)
            np.savez_compressed(os.path.join(HERE, f"transfer_{n}_m{mode}.npz"), T_g3=T, T_g1=T1)
    # deposit: interior, edges and out-of-bounds points, float32 positions
    n = 16
    npart = 300
    x = rng.uniform(-1.5, n + 0.5, npart).astype(np.float32)
    y = rng.uniform(-1.5, n + 0.5, npart).astype(np.float32)
    x[:8] = np.float32([1.0, n - 1.0, n - 1.0 - 2 ** -20, 0.999999, 1.5, 7.0, 7.25, 14.999])
    y[:8] = np.float32([1.0, 5.0, 5.0, 3.0, n - 1.0 - 2 ** -20, 7.0, 7.75, 1.0])
    m = rng.lognormal(0.0, 0.5, npart).astype(np.float32)
    for mode in (0, 1):
        rho = deposit_brute(n, mode, 40, x, y, m)
        np.savez_compressed(os.path.join(HERE, f"deposit_{n}_m{mode}.npz"), x=x, y=y, m=m, shift=40, rho=rho)
    # potential of a fixed random lattice
    for n, mode in ((16, 1), (32, 0), (32, 1)):
        rho_d = rng.uniform(0, 1, (n, n))
        T = transfer(n, mode, 1.0, 1e-3)
        np.savez_compressed(os.path.join(HERE, f"potential_{n}_m{mode}.npz"), rho=rho_d, phi=potential(n, T, rho_d))
    # 5-step all-FP64 trajectory
    n, npart, dt, g = 32, 64, 0.05, 1.0
    for mode in (0, 1):
        x = rng.uniform(8, 24, npart)
        y = rng.uniform(8, 24, npart)
        vx = rng.normal(0, 0.5, npart)
        vy = rng.normal(0, 0.5, npart)
        m = rng.uniform(0.5, 1.5, npart)
        T = transfer(n, mode, 1.0, 1e-3)
        x0, y0, vx0, vy0 = x.copy(), y.copy(), vx.copy(), vy.copy()
        for _ in range(5):
            rho, phi, fx, fy = step64(n, mode, dt, g, T, x, y, vx, vy, m)
        np.savez_compressed(os.path.join(HERE, f"traj64_{n}_m{mode}.npz"), x0=x0, y0=y0, vx0=vx0, vy0=vy0,
                            m=m, dt=dt, g=g, x=x, y=y, vx=vx, vy=vy, phi_last=phi, fx_last=fx, fy_last=fy)


if __name__ == "__main__":
    main()
