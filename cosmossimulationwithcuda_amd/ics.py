"""Seeded initial conditions for the benchmark and test workloads.

The reference draws its initial conditions from std::random_device
(CosmosCuda.cuh:936-937, 1083-1084), so they cannot be reproduced; these
generators are seeded so the GPU run and the CPU check see identical FP32
inputs.  Positions are in lattice cells, G is the force multiplier, and the
circular speed follows the lattice force law a = G M(<r) / r^2 of the 1/r
potential.  Returns float32 arrays (x, y, vx, vy, m).
"""
import os

import numpy as np


def _f32(*arrs):
    return tuple(np.ascontiguousarray(a, dtype=np.float32) for a in arrs)


def uniform_square(n, lo, hi, seed, mass=1.0):
    """n particles uniform in [lo, hi)^2 at rest, equal masses."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(lo, hi, n)
    y = rng.uniform(lo, hi, n)
    z = np.zeros(n)
    return _f32(x, y, z, z, np.full(n, mass))


def _disk_radii(rng, n, rd, rmax):
    out = np.empty(0)
    while out.size < n:
        k = int((n - out.size) * 1.1) + 64
        r = rd * rng.standard_gamma(2.0, k)        # surface density ~ exp(-r/rd): p(r) ~ r exp(-r/rd)
        out = np.concatenate([out, r[r < rmax]])
    return out[:n]


def _enclosed_fraction(r, rd, rmax):
    f = lambda s: 1.0 - (1.0 + s / rd) * np.exp(-s / rd)
    return f(r) / f(rmax)


def exponential_disk(n, cx, cy, rd, rmax, seed, gravity=1.0, masses="unit", dispersion=0.02,
                     bulk_vx=0.0, bulk_vy=0.0, rng=None):
    """Exponential disk: surface density ~ exp(-r/rd) truncated at rmax, uniform
    azimuth, circular speed from the enclosed mass, Gaussian velocity dispersion
    as a fraction of the local circular speed.  masses: "unit" or "lognormal"
    (mu = 0, sigma = 0.5)."""
    rng = rng if rng is not None else np.random.default_rng(seed)
    r = _disk_radii(rng, n, rd, rmax)
    a = rng.uniform(0.0, 2.0 * np.pi, n)
    m = np.ones(n) if masses == "unit" else rng.lognormal(0.0, 0.5, n)
    mtot = m.sum()
    vc = np.sqrt(gravity * mtot * _enclosed_fraction(r, rd, rmax) / np.maximum(r, 1e-3))
    ca, sa = np.cos(a), np.sin(a)
    x = cx + r * ca
    y = cy + r * sa
    vx = -vc * sa + dispersion * vc * rng.standard_normal(n) + bulk_vx
    vy = vc * ca + dispersion * vc * rng.standard_normal(n) + bulk_vy
    return _f32(x, y, vx, vy, m)


def circular_speed(mtot, rd, rmax, r, gravity=1.0):
    return float(np.sqrt(gravity * mtot * _enclosed_fraction(np.float64(r), rd, rmax) / r))


def colliding_disks(n, seed, n_lattice=2048, rd=96.0, centres=((700.0, 1024.0), (1348.0, 1024.0)),
                    approach=0.3, gravity=1.0):
    """Two exponential disks (n/2 particles each, unit masses, truncated at 5 rd)
    approaching along x at +/- approach * v_c(rd)."""
    rng = np.random.default_rng(seed)
    half = n // 2
    rmax = 5.0 * rd
    parts = []
    for k, (cx, cy) in enumerate(centres):
        nk = half if k == 0 else n - half
        v = approach * circular_speed(nk, rd, rmax, rd, gravity)
        parts.append(exponential_disk(nk, cx, cy, rd, rmax, None, gravity, "unit", 0.02,
                                      bulk_vx=v if k == 0 else -v, rng=rng))
    return tuple(np.concatenate([p[i] for p in parts]) for i in range(5))


def single_disk(n, seed, gravity=1.0):
    """2048^2 single disk: rd = 128 cells, truncated at 900, centre (1024, 1024),
    lognormal masses, 2% dispersion."""
    return exponential_disk(n, 1024.0, 1024.0, 128.0, 900.0, seed, gravity, "lognormal", 0.02)


def plummer_binary(n, seed, n_lattice=4096, a=64.0, separation=800.0, gravity=1.0):
    """Two Plummer spheres (projected onto the plane), scale radius a, unit
    masses, on a circular mutual orbit at the given separation."""
    rng = np.random.default_rng(seed)
    c = n_lattice / 2.0
    half = n // 2
    mtot = float(n)
    vorb = np.sqrt(gravity * (mtot / 2.0) / (2.0 * separation))   # each body: M/2 at distance sep, radius sep/2
    out = []
    for k, sgn in enumerate((-1.0, 1.0)):
        nk = half if k == 0 else n - half
        u = rng.uniform(0.0, 0.99, nk)
        r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        r = np.minimum(r, 6.0 * a)
        ct = rng.uniform(-1.0, 1.0, nk)
        rp = r * np.sqrt(1.0 - ct ** 2)            # projected radius
        ph = rng.uniform(0.0, 2.0 * np.pi, nk)
        x = c + sgn * separation / 2.0 + rp * np.cos(ph)
        y = c + rp * np.sin(ph)
        vx = np.zeros(nk)
        vy = np.full(nk, sgn * vorb)
        out.append(_f32(x, y, vx, vy, np.ones(nk)))
    return tuple(np.concatenate([p[i] for p in out]) for i in range(5))


# ----- serving ensemble -------------------------------------------------------------
# Requests of the serving workload are disks centred on their lattice, in
# lattice-scaled units: positions are s * (template), s = n_lattice / 16, and
# with unit masses and G = 0.25 * s / n the circular speed does not depend on
# the particle count n or the lattice (v = 0.5 * template speed, <= ~0.3
# cells per time unit).  Templates hold i.i.d. draws, so a contiguous slice at
# a seeded offset is itself a seeded sample of the disk.

def disk_template(kind, n, seed, dispersion=0.02):
    """Unit-scale disk of n particles about the origin, circular speeds for
    total mass 1 and G = 1 (1/r force law).  kind: "exponential" (r_d = 1,
    truncated at 5) or "uniform" (radius 4)."""
    rng = np.random.default_rng(seed)
    if kind == "exponential":
        r = _disk_radii(rng, n, 1.0, 5.0)
        frac = _enclosed_fraction(r, 1.0, 5.0)
    elif kind == "uniform":
        r = 4.0 * np.sqrt(rng.uniform(0.0, 1.0, n))
        frac = (r / 4.0) ** 2
    else:
        raise ValueError(kind)
    a = rng.uniform(0.0, 2.0 * np.pi, n)
    vc = np.sqrt(frac / np.maximum(r, 1e-3))
    ca, sa = np.cos(a), np.sin(a)
    vx = -vc * sa + dispersion * vc * rng.standard_normal(n)
    vy = vc * ca + dispersion * vc * rng.standard_normal(n)
    return _f32(r * ca, r * sa, vx, vy)


def ensemble_request(template, n_lattice, n, seed):
    """(x, y, vx, vy, m, gravity) of one serving request: n particles sliced
    from a disk template at a seeded offset, centred on an n_lattice lattice
    (with a seeded sub-cell jitter of the centre)."""
    tx, ty, tvx, tvy = template
    rng = np.random.default_rng(seed)
    off = int(rng.integers(0, tx.size - n + 1))
    s = np.float32(n_lattice / 16.0)
    cx, cy = np.float32(n_lattice / 2.0 + rng.uniform(-0.5, 0.5)), np.float32(n_lattice / 2.0 + rng.uniform(-0.5, 0.5))
    sl = slice(off, off + n)
    half = np.float32(0.5)
    return (cx + s * tx[sl], cy + s * ty[sl], half * tvx[sl], half * tvy[sl], np.ones(n, np.float32),
            0.25 * float(s) / n)


# ----- parallel generation for very large sets ---------------------------------------

def _disk_chunk(args):
    (names, off, cnt, cx, cy, rd, rmax, mtot, bulk_vx, seed, gravity) = args
    from multiprocessing import shared_memory
    rng = np.random.default_rng(seed)
    r = _disk_radii(rng, cnt, rd, rmax)
    a = rng.uniform(0.0, 2.0 * np.pi, cnt)
    vc = np.sqrt(gravity * mtot * _enclosed_fraction(r, rd, rmax) / np.maximum(r, 1e-3))
    ca, sa = np.cos(a), np.sin(a)
    vals = (cx + r * ca, cy + r * sa,
            -vc * sa + 0.02 * vc * rng.standard_normal(cnt) + bulk_vx,
            vc * ca + 0.02 * vc * rng.standard_normal(cnt))
    shms = [shared_memory.SharedMemory(name=nm) for nm in names]
    try:
        for shm, v in zip(shms[:4], vals):
            np.ndarray((off + cnt,), np.float32, buffer=shm.buf)[off:off + cnt] = v
        np.ndarray((off + cnt,), np.float32, buffer=shms[4].buf)[off:off + cnt] = 1.0
    finally:
        for shm in shms:
            shm.close()
    return cnt


def colliding_disks_parallel(n, seed, workers=None, chunk=4_000_000, n_lattice=2048, rd=96.0,
                             centres=((700.0, 1024.0), (1348.0, 1024.0)), approach=0.3, gravity=1.0):
    """colliding_disks' law for very large n, generated in chunks by worker
    processes (seeded per chunk, so the set depends only on n, seed and
    chunk; it is a different draw from colliding_disks(n, seed)).  Returns
    float32 arrays (x, y, vx, vy, m)."""
    import multiprocessing as mp
    from multiprocessing import shared_memory
    workers = workers or min(16, os.cpu_count() or 1)
    half = n // 2
    rmax = 5.0 * rd
    shms = [shared_memory.SharedMemory(create=True, size=max(4 * n, 4)) for _ in range(5)]
    try:
        names = [s.name for s in shms]
        jobs = []
        off = 0
        for k, (cx, cy) in enumerate(centres):
            nk = half if k == 0 else n - half
            v = approach * circular_speed(nk, rd, rmax, rd, gravity)
            for c0 in range(0, nk, chunk):
                cnt = min(chunk, nk - c0)
                jobs.append((names, off + c0, cnt, cx, cy, rd, rmax, float(nk), v if k == 0 else -v,
                             (seed, k, c0 // chunk), gravity))
            off += nk
        with mp.get_context("spawn").Pool(workers) as pool:
            pool.map(_disk_chunk, jobs)
        return tuple(np.ndarray((n,), np.float32, buffer=s.buf).copy() for s in shms)
    finally:
        for s in shms:
            s.close()
            s.unlink()
