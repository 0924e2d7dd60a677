"""Headline benchmark: particle-steps/s of the PM gravity step on one MI355X.

Workload (default): 2048^2 lattice, 100,000,000 particles, CIC, two colliding
exponential disks (r_d = 96 cells, 5e7 unit-mass particles each, centres
(700,1024) and (1348,1024), +/-0.3 v_c(r_d) approach), dt = 0.002.  One step =
deposit + FFT field solve + force gather / kick-drift over the full particle
set, inputs resident in HBM.  W warm-up steps, then K timed steps bracketed by
device synchronisation.  Per-kernel durations come from HIP events recorded on
the simulation's own stream inside the timed region.

Prints ONE JSON line (rank 0, N=1).  --gpus N>1 exits: this engine runs one
simulation per GPU and has no multi-GPU path.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s @100M CIC 2048²; ensemble steps/s + p99 latency vs plain"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (lattice, particles, mode, dt, generator)
    "colliding_disks_2048_100M_cic": (2048, 100_000_000, 1, 0.002, "colliding"),
    "colliding_disks_2048_100M_ngp": (2048, 100_000_000, 0, 0.002, "colliding"),
    "single_disk_2048_20M_cic": (2048, 20_000_000, 1, 0.005, "single"),
    # full one-GPU scale: the same two disks with 5e8 particles each
    "colliding_disks_2048_1B_cic": (2048, 1_000_000_000, 1, 0.002, "colliding_parallel"),
    "colliding_disks_2048_1B_ngp": (2048, 1_000_000_000, 0, 0.002, "colliding_parallel"),
}


def algorithmic_bytes(kernel, n, np_):
    """Algorithmic HBM bytes per launch: particle streams the kernel must move
    plus the lattice it must read or write once (implementation extras such as
    atomics' read-modify-write or re-reads are not counted)."""
    nk = n // 2 + 1
    return {
        # within pm_step only the bin counts of the next re-sort are cleared
        # by this timer (the row pass zeroes the next int64 lattice)
        "clear": 4 * ((n // (32 if n <= 2048 else 64)) ** 2 + 1),
        "deposit": 12 * np_ + 8 * n * n,          # read x, y, m; write int64 lattice
        # read int64 lattice, write half spectrum, zero the next int64 lattice
        "row_r2c": 8 * n * n + 8 * nk * n + 8 * n * n,
        "col_conv": 8 * nk * n + 4 * nk * n + 16 * nk * n,   # read H, read T, write H and Hfy
        "row_c2r": 16 * nk * n + 4 * n * n + 8 * n * n,      # read H, Hfy; write Phi and force
        "edge": 8 * 8 * n,
        "stencil": 4 * n * n + 8 * n * n,
        "push": 32 * np_ + 8 * n * n,             # read/write x, y, vx, vy; read force lattice
        "sort": 0,                                # re-binning is implementation traffic, not algorithmic
        # fused push + next deposit: read x, y, vx, vy, m; write x, y, vx, vy; read force; write int64 lattice
        "push_deposit": 36 * np_ + 16 * n * n,
        # the same step with the re-sort fused in (moving m and the load index is sort traffic)
        "push_deposit_sort": 36 * np_ + 16 * n * n,
        # the same step kernel on the step that also counts the next re-sort's bins
        "push_deposit_count": 36 * np_ + 16 * n * n,
    }[kernel]


def make_ics(gen, np_, seed):
    from cosmossimulationwithcuda_amd import ics
    if gen == "colliding":
        return ics.colliding_disks(np_, seed)
    if gen == "colliding_parallel":
        return ics.colliding_disks_parallel(np_, seed)
    return ics.single_disk(np_, seed)


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor()


def cpu_baseline(n, mode, dt, arrs, steps, threads, shift):
    """The FP32 CPU step (oracle/pm_cpu32.c: int64 deposit into per-thread
    lattices allocated once, FP32 FFT convolution, the reference's FP32
    stencil and push) on the same particles, timed on this host's cores."""
    from oracle import oracle as orc
    orc.set_threads(threads)
    x, y, vx, vy, m = arrs
    s = orc.Cpu32Stepper(n, mode, np.float32(dt), 1.0, 1.0, 1e-3, 0, shift, x, y, vx, vy, m)
    s.step(1)                         # warm-up (page-in, thread start)
    t0 = time.perf_counter()
    s.step(steps)
    el = time.perf_counter() - t0
    return len(x) * steps / el, el, orc.num_threads()


def device_copy_rate(nbytes=2 << 30, reps=10):
    """Bytes moved per second (read + write) by the HIP runtime's own
    device-to-device copy of nbytes, HIP events around reps copies: what a
    plain streaming kernel of the same device reaches, for scale."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    a, b, e0, e1 = vp(), vp(), vp(), vp()
    assert hip.hipMalloc(ctypes.byref(a), sz(nbytes)) == 0 and hip.hipMalloc(ctypes.byref(b), sz(nbytes)) == 0
    assert hip.hipMemset(a, 1, sz(nbytes)) == 0
    assert hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0
    for _ in range(2):
        assert hip.hipMemcpyAsync(b, a, sz(nbytes), 3, vp(0)) == 0          # hipMemcpyDeviceToDevice
    hip.hipEventRecord(e0, vp(0))
    for _ in range(reps):
        hip.hipMemcpyAsync(b, a, sz(nbytes), 3, vp(0))
    hip.hipEventRecord(e1, vp(0))
    hip.hipEventSynchronize(e1)
    ms = ctypes.c_float()
    hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
    for x in (e0, e1):
        hip.hipEventDestroy(x)
    hip.hipFree(a)
    hip.hipFree(b)
    return 2 * nbytes * reps / (ms.value / 1e3) / 1e9


def copy_probe(nbytes):
    """One device-to-device hipMemcpy of nbytes (a copy kernel of known size)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 0, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0      # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"# copy probe: {nbytes} bytes device to device", file=sys.stderr, flush=True)


# Precision variant: 4096^2 lattice, 1e8 particles, Plummer binary (a = 64
# cells, 5e7 each, circular orbit at separation 800), CIC.  dt = 5e-4: the
# spheres start cold and collapse; their dynamical time sqrt(a^3 / (G M)) is
# 0.072, and at dt = 2e-3 the collapsed cores move several cells per step.
PRECISION = {"lattice": 4096, "particles": 100_000_000, "dt": 0.0005, "samples": (0, 5, 10, 20), "timed_steps": 20}


def precision_report(seed):
    """FP32 FFT vs split-BF16 matrix-core FFT on the same workload: throughput,
    and at sample steps the potential / force / total-energy error of each
    against the FP64 oracle field of that step's own deposited lattice."""
    from cosmossimulationwithcuda_amd import ics, pm
    from oracle import oracle as orc
    n, npart, dt = PRECISION["lattice"], PRECISION["particles"], PRECISION["dt"]
    t = time.perf_counter()
    arrs = ics.plummer_binary(npart, seed, n_lattice=n)
    print(f"# precision: ics {npart} in {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
    T = orc.transfer(n, pm.CIC, 1.0, 1e-3)
    out = {"workload": f"plummer_binary_{n}_{npart // 1_000_000}M_cic", "lattice": n, "particles": npart, "dt": dt}
    for name, prec in (("fp32", pm.FFT_FP32), ("bf16x3", pm.FFT_BF16X3), ("bf16x6", pm.FFT_BF16X6)):
        cfg = pm.default_config(n_lattice=n, n_particles=npart, mode=pm.CIC, dt=dt, gravity=1.0, green_origin=1.0,
                                deconv_eps=1e-3, precision=prec)
        rows = []
        with pm.Simulation(cfg) as s:
            s.load(*arrs)
            done = 0
            for k in PRECISION["samples"]:
                if k == 0:
                    s.stage(pm.STAGE_DEPOSIT)
                s.step(k - done)
                done = k
                s.stage(pm.STAGE_FIELD)                  # field of the current lattice
                rho = s.read_lattice(pm.LATTICE_MASS)
                phi = s.read_lattice(pm.LATTICE_POTENTIAL).astype(np.float64)
                force = s.read_lattice(pm.LATTICE_FORCE).astype(np.float64)
                shift = s.info().mass_shift
                ke = s.diagnostics()["kinetic"]
                phi64 = orc.potential(n, T, shift, rho)
                fx64, fy64 = orc.gradient64(n, 1.0, phi64)
                mass = rho.astype(np.float64) * 2.0 ** -shift
                e64 = ke + 0.5 * float(np.sum(mass * phi64))
                e = ke + 0.5 * float(np.sum(mass * phi))
                rows.append({"step": k,
                             "potential_rel_l2": float(np.linalg.norm(phi - phi64) / np.linalg.norm(phi64)),
                             "force_rel_l2": float(np.sqrt(np.sum((force[..., 0] - fx64) ** 2 + (force[..., 1] - fy64) ** 2)
                                                           / np.sum(fx64 ** 2 + fy64 ** 2))),
                             "energy_rel_err": abs(e - e64) / abs(e64)})
                print(f"# precision {name} step {k}: {rows[-1]}", file=sys.stderr, flush=True)
            s.profile(True)
            s.profile_reset()
            s.sync()
            t0 = time.perf_counter()
            s.step(PRECISION["timed_steps"])
            s.sync()
            el = time.perf_counter() - t0
            prof = s.profile_read()
        fft_ms = sum(prof[k][0] / max(prof[k][1], 1) for k in ("row_r2c", "col_conv", "row_c2r"))
        out[name] = {"particle_steps_per_s": npart * PRECISION["timed_steps"] / el,
                     "ms_per_step": el / PRECISION["timed_steps"] * 1e3, "fft_ms_per_step": round(fft_ms, 4),
                     "kernels_ms": {k: [round(v[0] / v[1], 4), v[1]] for k, v in prof.items() if v[1]},
                     "error_vs_fp64": rows}
    return out


def field_report(sizes, reps, seed):
    """The three FFT passes of the field solve (stage FIELD: row R2C, column
    pass with the transfer function and the y derivative, row C2R with the
    edge bands) per lattice size, one simulation, default plan: average
    duration from HIP events and achieved GB/s of the algorithmic bytes
    (the R2C pass of the stage API does not zero a next lattice)."""
    from cosmossimulationwithcuda_amd import ics, pm
    out = {}
    for n in sizes:
        np_ = 1_000_000
        arrs = ics.uniform_square(np_, n // 4, 3 * n // 4, seed=seed)
        cfg = pm.default_config(n_lattice=n, n_particles=np_, mode=pm.CIC, dt=0.001, gravity=1.0, green_origin=1.0,
                                deconv_eps=1e-3)
        with pm.Simulation(cfg) as s:
            s.load(*arrs)
            s.stage(pm.STAGE_DEPOSIT)
            s.stage(pm.STAGE_FIELD)
            s.profile(True)
            s.profile_reset()
            for _ in range(reps):
                s.stage(pm.STAGE_FIELD)
            s.sync()
            prof = s.profile_read()
            plan = {k: s.get_param(k) for k in ("fft_radix", "fft_it")}
        nk = n // 2 + 1
        row = {"plan": plan}
        for k, gb in (("row_r2c", (8 * n * n + 8 * nk * n) / 1e9), ("col_conv", algorithmic_bytes("col_conv", n, 0) / 1e9),
                      ("row_c2r", algorithmic_bytes("row_c2r", n, 0) / 1e9)):
            ms, cnt = prof[k]
            us = ms / cnt * 1e3
            row[k] = {"avg_us": round(us, 2), "algorithmic_MB": round(gb * 1e3, 2), "GB_s": round(gb / (us * 1e-6), 1)}
        row["vendor_fft_reference_us"] = vendor_fft_reference(n, reps)
        out[str(n)] = row
        print(f"# field {n}: {row}", file=sys.stderr, flush=True)
    return out


def vendor_fft_reference(n, reps):
    """The ROCm FFT library (hipFFT over rocFFT) on the same device, for scale:
    a batched 1D C2C of N/2 rows of length N (the work of one row pass: two
    real rows per complex transform), and the plain 2D R2C and C2R of an N^2
    lattice.  Average us per call, HIP events on the null stream; None if the
    library is absent."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        fft = ctypes.CDLL("libhipfft.so")
    except OSError:
        return None
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    nn = n * n
    a, b = vp(), vp()
    assert hip.hipMalloc(ctypes.byref(a), sz(nn * 16)) == 0 and hip.hipMalloc(ctypes.byref(b), sz(nn * 16)) == 0
    assert hip.hipMemset(a, 0, sz(nn * 16)) == 0 and hip.hipMemset(b, 0, sz(nn * 16)) == 0
    ev0, ev1 = vp(), vp()
    assert hip.hipEventCreate(ctypes.byref(ev0)) == 0 and hip.hipEventCreate(ctypes.byref(ev1)) == 0

    def timed(call):
        for _ in range(3):
            assert call() == 0
        hip.hipEventRecord(ev0, vp(0))
        for _ in range(reps):
            assert call() == 0
        hip.hipEventRecord(ev1, vp(0))
        hip.hipEventSynchronize(ev1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), ev0, ev1)
        return round(ms.value / reps * 1e3, 2)

    C2C, R2C, C2R, FWD = 0x29, 0x2a, 0x2c, -1
    for f in (fft.hipfftExecC2C, fft.hipfftExecR2C, fft.hipfftExecC2R, fft.hipfftDestroy):
        f.restype = ctypes.c_int
    fft.hipfftExecC2C.argtypes = [vp, vp, vp, ctypes.c_int]
    fft.hipfftExecR2C.argtypes = fft.hipfftExecC2R.argtypes = [vp, vp, vp]
    fft.hipfftDestroy.argtypes = [vp]
    out = {}
    plans = [vp(), vp(), vp()]                   # hipfftHandle is a pointer on ROCm
    length = (ctypes.c_int * 1)(n)
    assert fft.hipfftPlanMany(ctypes.byref(plans[0]), 1, length, None, 1, n, None, 1, n, C2C, n // 2) == 0
    out["c2c_1d_rows_half"] = timed(lambda: fft.hipfftExecC2C(plans[0], a, b, FWD))
    assert fft.hipfftPlan2d(ctypes.byref(plans[1]), n, n, R2C) == 0
    out["r2c_2d"] = timed(lambda: fft.hipfftExecR2C(plans[1], a, b))
    assert fft.hipfftPlan2d(ctypes.byref(plans[2]), n, n, C2R) == 0
    out["c2r_2d"] = timed(lambda: fft.hipfftExecC2R(plans[2], b, a))
    for q in plans:
        fft.hipfftDestroy(q)
    hip.hipEventDestroy(ev0)
    hip.hipEventDestroy(ev1)
    hip.hipFree(a)
    hip.hipFree(b)
    return out


# Serving ensemble: independent requests arriving as a Poisson stream.
SERVING = {"requests": 512, "rate_per_s": 40.0, "lattices": (256, 512, 1024, 2048), "particles": (1e4, 1e7),
           "steps": 200, "budget_bytes": 200 << 30, "dt": 1.0, "template": 10_000_000,
           # pinned host staging for the submitted arrays: the 512 requests hold 13.7 GB
           "pinned_mib": 20480}


def serving_workload(nreq, seed):
    """Seeded request list: lattice uniform over the sizes, particle count
    log-uniform, disk kind (exponential / uniform) uniform, Poisson arrivals."""
    rng = np.random.default_rng(seed)
    lo, hi = SERVING["particles"]
    lat = rng.choice(SERVING["lattices"], nreq)
    npart = np.exp(rng.uniform(np.log(lo), np.log(hi), nreq)).astype(np.int64)
    kind = rng.integers(0, 2, nreq)
    arrive = np.cumsum(rng.exponential(1.0 / SERVING["rate_per_s"], nreq))
    seeds = rng.integers(0, 2 ** 31, nreq)
    return [dict(n_lattice=int(a), n=int(b), kind=int(c), t=float(d), seed=int(e))
            for a, b, c, d, e in zip(lat, npart, kind, arrive, seeds)]


def serving_run(pm, reqs, arrays, mode, poisson, use_graphs=1):
    """Submits the requests (at their arrival times, or all at once) with their
    pre-generated arrays, waits for every result, returns the server
    statistics."""
    import queue
    import threading
    t = time.perf_counter()
    srv = pm.Server(pm.default_server_config(mode=mode, hbm_budget_bytes=SERVING["budget_bytes"],
                                             pinned_staging_mib=SERVING["pinned_mib"], use_graphs=use_graphs))
    t_create = time.perf_counter() - t
    q = queue.Queue()
    errors = []
    submitted = []

    def submitter():
        try:
            t0 = time.perf_counter()
            for r in reqs:
                if poisson:
                    d = t0 + r["t"] - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                x, y, vx, vy, m, g = arrays[len(submitted)]
                cfg = pm.default_config(n_lattice=r["n_lattice"], n_particles=r["n"], mode=pm.CIC,
                                        dt=SERVING["dt"], gravity=g, green_origin=1.0, deconv_eps=1e-3)
                q.put(srv.submit(cfg, x, y, vx, vy, m, SERVING["steps"]))
                submitted.append(r)
        except Exception as e:          # reported by the waiting side
            errors.append(e)
        q.put(None)

    th = threading.Thread(target=submitter)
    th.start()
    done = 0
    while True:
        rid = q.get()
        if rid is None:
            break
        srv.wait(rid, copy=False)
        done += 1
        if done % 128 == 0:
            print(f"#   {done} requests done", file=sys.stderr, flush=True)
    th.join()
    st = srv.stats()
    srv.close()
    if errors:
        raise errors[0]
    wall = st["last_done_s"] - st["first_submit_s"]
    return {"wall_s": round(wall, 3),
            "steps_per_s": st["steps"] / wall,
            "particle_steps_per_s": st["particle_steps"] / wall,
            "step_latency_p50_ms": round(st["step_latency_p50_ms"], 4),
            "step_latency_p99_ms": round(st["step_latency_p99_ms"], 4),
            "request_latency_p50_ms": round(st["request_latency_p50_ms"], 2),
            "request_latency_p99_ms": round(st["request_latency_p99_ms"], 2),
            "peak_GB": round(st["peak_bytes"] / 1e9, 2), "rounds": st["rounds"],
            "server_create_s": round(t_create, 3), "pinned_staging_GiB": SERVING["pinned_mib"] / 1024,
            "graph_launches": st["graph_launches"], "batched_rounds": st["batched_launch_rounds"],
            "pool_hits": st["pool_hits"], "graph_captures": st["graph_captures"], "failed": st["failed"],
            "worker_s": {k: round(st[k], 3) for k in ("admit_s", "launch_s", "readback_s", "wait_s", "create_s",
                                                     "upload_s")}}


def serving_naive_loop(pm, reqs, arrays):
    """The literal one-at-a-time loop a caller writes without the server: for
    each request create a simulation, load, step, read back, destroy."""
    wall = 0.0
    steps = psteps = 0
    for r, (x, y, vx, vy, m, g) in zip(reqs, arrays):
        cfg = pm.default_config(n_lattice=r["n_lattice"], n_particles=r["n"], mode=pm.CIC, dt=SERVING["dt"],
                                gravity=g, green_origin=1.0, deconv_eps=1e-3)
        t0 = time.perf_counter()
        with pm.Simulation(cfg) as s:
            s.load(x, y, vx, vy, m)
            s.step(SERVING["steps"])
            s.read()
        wall += time.perf_counter() - t0
        steps += SERVING["steps"]
        psteps += SERVING["steps"] * r["n"]
    return {"wall_s": round(wall, 3), "steps_per_s": steps / wall, "particle_steps_per_s": psteps / wall,
            "note": "time of create + load + steps + read-back + destroy per request, summed; arrays generated "
                    "before timing, as for the server modes"}


def shift_of(m):
    """The fixed-point shift the engine picks for these masses: floor(61 -
    log2(sum |m|)), clamped to [0, 62]."""
    s = float(np.sum(np.abs(m.astype(np.float64))))
    return 40 if not s > 0 else max(0, min(62, int(np.floor(61.0 - np.log2(s)))))


def serving_cpu_baseline(ics, reqs, templates, steps=5):
    """The one-at-a-time loop on the host cores with the FP32 CPU step
    (oracle/pm_cpu32.c) over a subset of the requests, a few steps each (a
    bounded sample, set-up included)."""
    from oracle import oracle as orc
    threads = os.cpu_count()
    orc.set_threads(threads)
    el = 0.0
    psteps = 0
    for r in reqs:
        x, y, vx, vy, m, g = ics.ensemble_request(templates[r["kind"]], r["n_lattice"], r["n"], r["seed"])
        t0 = time.perf_counter()
        st = orc.Cpu32Stepper(r["n_lattice"], 1, np.float32(SERVING["dt"]), np.float32(g), 1.0, 1e-3, 0,
                              shift_of(m), x, y, vx, vy, m)
        st.step(steps)
        el += time.perf_counter() - t0
        psteps += steps * r["n"]
    return {"value": psteps / el, "unit": "particle-steps/s", "cores": orc.num_threads(), "kind": "port",
            "sample": f"first {len(reqs)} requests of the ensemble, {steps} steps each (set-up included), "
                      f"{el:.1f}s; oracle/pm_cpu32.c (FP32 CPU step), host {cpu_model()}"}


def serving_bench(nreq, seed, modes=("plain", "batched"), arrivals=("saturated", "poisson"), use_graphs=1):
    from cosmossimulationwithcuda_amd import ics, pm
    t = time.perf_counter()
    templates = [ics.disk_template(k, SERVING["template"], seed + i) for i, k in enumerate(("exponential", "uniform"))]
    reqs = serving_workload(nreq, seed)
    # Every request's arrays are generated once, before any timed mode: all
    # modes then time the same work (generating 6.9e8 particles takes ~4 s on
    # one host core, as long as the batched server's whole run).
    arrays = [ics.ensemble_request(templates[r["kind"]], r["n_lattice"], r["n"], r["seed"]) for r in reqs]
    print(f"# serving: {nreq} requests, {sum(r['n'] for r in reqs)} particles total, arrays in "
          f"{time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
    out = {"workload": f"ensemble_{nreq}req_poisson{SERVING['rate_per_s']:g}_{SERVING['steps']}steps_cic",
           "requests": nreq, "steps_per_request": SERVING["steps"],
           "lattices": list(SERVING["lattices"]), "particles_loguniform": list(SERVING["particles"]),
           "budget_GB": SERVING["budget_bytes"] / 2 ** 30, "use_graphs": use_graphs}
    # One untimed pass over the first 32 requests in each server mode: the
    # first server run of a process pays one-time costs (code-object loading,
    # first device allocations, graph instantiation; ~0.7 s measured) that
    # would otherwise fall on whichever mode is timed first.
    for name, mode in (("plain", pm.SERVE_PLAIN), ("batched", pm.SERVE_BATCHED)):
        if name in modes:
            print(f"# serving warm-up {name}", file=sys.stderr, flush=True)
            serving_run(pm, reqs[:32], arrays[:32], mode, False, use_graphs)
    for tag, poisson in (("saturated", False), ("poisson", True)):
        if tag not in arrivals:
            continue
        for name, mode in (("plain", pm.SERVE_PLAIN), ("batched", pm.SERVE_BATCHED)):
            if name not in modes:
                continue
            print(f"# serving {tag} {name}", file=sys.stderr, flush=True)
            out[f"{tag}_{name}"] = serving_run(pm, reqs, arrays, mode, poisson, use_graphs)
        if "plain" in modes and "batched" in modes:
            p, b = out[f"{tag}_plain"], out[f"{tag}_batched"]
            out[f"{tag}_speedup_steps_per_s"] = round(b["steps_per_s"] / p["steps_per_s"], 3)
    if "cpu" in modes:
        out["cpu_baseline"] = serving_cpu_baseline(ics, reqs[:16], templates)
    if "saturated" in arrivals and "naive" in modes:
        print("# serving naive per-request loop", file=sys.stderr, flush=True)
        out["naive_loop"] = serving_naive_loop(pm, reqs, arrays)
        if "saturated_batched" in out:
            out["saturated_speedup_vs_naive_loop"] = round(
                out["saturated_batched"]["steps_per_s"] / out["naive_loop"]["steps_per_s"], 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="colliding_disks_2048_100M_cic", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-steps", type=int, default=60,
                    help="timed FP32 CPU steps for cpu_baseline at 1e8 particles, scaled by the count (0 = skip)")
    ap.add_argument("--breakdown-steps", type=int, default=16,
                    help="steps of the profiled pass after the timed region that times the non-step kernels")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores")
    ap.add_argument("--window-steps", type=int, default=180,
                    help="after the timed region: reload and time steps 20..20+this (0 = skip)")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the run, copy 1 GiB device-to-device (hipMemcpy) as the calibration "
                         "kernel of known size for rocprofv3 FETCH_SIZE / WRITE_SIZE passes")
    ap.add_argument("--serving", choices=("on", "off", "only"), default="on",
                    help="serving ensemble leg (plain vs batched server, saturated and Poisson arrivals)")
    ap.add_argument("--precision-report", action="store_true",
                    help="only the 4096^2 Plummer-binary FP32 vs split-BF16 (bf16x3, bf16x6) FFT comparison "
                         "(error vs FP64 per step)")
    ap.add_argument("--field-report", default="",
                    help="comma-separated lattice sizes: time the FFT passes of the field solve at each and exit")
    ap.add_argument("--field-reps", type=int, default=20)
    ap.add_argument("--serving-requests", type=int, default=SERVING["requests"])
    ap.add_argument("--serving-modes", default="plain,batched,naive,cpu")
    ap.add_argument("--serving-arrivals", default="saturated,poisson")
    ap.add_argument("--serving-graphs", type=int, default=1, help="replay rounds from captured HIP graphs")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs each simulation on one GPU; multi-GPU runs are not supported "
              f"(asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)

    from cosmossimulationwithcuda_amd import pm

    if args.precision_report:
        print(json.dumps({"precision": precision_report(args.seed)}), flush=True)
        return
    if args.field_report:
        print(json.dumps({"field": field_report([int(v) for v in args.field_report.split(",")], args.field_reps,
                                                args.seed)}), flush=True)
        return
    if args.serving == "only":
        print(json.dumps({"serving": serving_bench(args.serving_requests, args.seed, args.serving_modes.split(","),
                                   args.serving_arrivals.split(","), args.serving_graphs)}), flush=True)
        return

    n, np_, mode, dt, gen = WORKLOADS[args.workload]
    t_ic = time.perf_counter()
    arrs = make_ics(gen, np_, args.seed)
    t_ic = time.perf_counter() - t_ic
    print(f"# ics {args.workload}: {np_} particles in {t_ic:.1f}s", file=sys.stderr, flush=True)

    cfg = pm.default_config(n_lattice=n, n_particles=np_, mode=mode, dt=dt, gravity=1.0, green_origin=1.0,
                            deconv_eps=1e-3)
    sim = pm.Simulation(cfg)
    sim.load(*arrs)
    sim.step(args.warmup)
    sim.sync()
    print(f"# warm-up {args.warmup} steps done", file=sys.stderr, flush=True)

    # Timed region: HIP events only around the step kernels (the roofline's
    # kernel); timing every launch would add ~30 us of event records per step.
    step_kernels = ("push_deposit", "push_deposit_sort", "push_deposit_count", "sort")
    sim.set_param("profile_mask", sum(1 << pm.KERNELS.index(k) for k in step_kernels))
    sim.profile(True)
    sim.profile_reset()
    sim.sync()
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.sync()
    el = time.perf_counter() - t0
    prof = sim.profile_read()
    print(f"# timed {args.steps} steps in {el:.3f}s", file=sys.stderr, flush=True)
    # The other kernels' durations, for the breakdown: a separate profiled pass.
    sim.set_param("profile_mask", -1)
    sim.profile_reset()
    sim.step(args.breakdown_steps)
    sim.sync()
    prof_rest = sim.profile_read()
    sim.profile(False)
    prof = {k: (prof[k] if k in step_kernels else prof_rest[k]) for k in prof}
    diag = sim.diagnostics()
    shift = sim.info().mass_shift

    kernels = {}
    for name, (ms, cnt) in prof.items():
        if cnt == 0:
            continue
        avg_ms = ms / cnt
        gb = algorithmic_bytes(name, n, np_) / 1e9
        kernels[name] = {"avg_ms": round(avg_ms, 5), "launches": cnt, "algorithmic_GB": round(gb, 4),
                         "GB_s": round(gb / (avg_ms / 1e3), 1) if avg_ms > 0 else None}
    # the steady step kernel; with a re-sort every 2 steps every step is either
    # the bin-counting or the re-sorting variant, and the counting one is the
    # step kernel proper
    dom = next((k for k in ("push_deposit", "push_deposit_count") if k in kernels),
               max(("deposit", "push"), key=lambda k: kernels.get(k, {"avg_ms": 0})["avg_ms"]))
    kd = kernels[dom]
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get(args.workload, {}).get(dom)
        except Exception:
            traffic = None
    # Amortised over the re-sort cycle: every launch of the particle step
    # (steady, bin-counting and re-sorting variants, plus any separate sort
    # pass, whose bytes are not algorithmic) -> summed bytes / summed time.
    am_bytes = am_ms = 0.0
    for k in step_kernels:
        ms, cnt = prof.get(k, (0.0, 0))
        if cnt:
            am_bytes += algorithmic_bytes(k, n, np_) * cnt
            am_ms += ms
    frac_am = round(am_bytes / (am_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if am_ms > 0 else None
    copy_gbs = round(device_copy_rate(), 1)
    roofline = {"bound": "hbm", "achieved": kd["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(kd["GB_s"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": "profiles/traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel, "
                                  "committed; not measured in this run)" if traffic is not None else None,
                "kernel": dom, "frac_amortised": frac_am,
                "device_copy_GB_s": copy_gbs,
                "device_copy_note": "hipMemcpy device to device of 2 GiB (read + write bytes per second) timed in "
                                    "the same run: the rate of the runtime's own streaming copy kernel",
                "frac_amortised_kernels": [k for k in step_kernels if prof.get(k, (0, 0))[1]]}

    value = np_ * args.steps / el
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "lattice": n, "particles": np_,
                   "mode": "cic" if mode == 1 else "ngp", "dt": dt, "force_mode": "spectral"},
        "roofline": roofline,
        "kernels": kernels,
        "kernels_note": "push_deposit* timed with HIP events in the timed region; the other kernels in a "
                        f"separate {args.breakdown_steps}-step profiled pass after it",
        "diagnostics": {k: diag[k] for k in ("mass", "kinetic", "potential", "nonfinite", "in_bounds")},
    }
    # Step time over the configuration's own window (steps 20-200 after a
    # fresh load): the disks collapse as they approach, so the rate may drift.
    if args.window_steps > 0:
        sim.load(*arrs)
        sim.step(20)
        sim.sync()
        t0 = time.perf_counter()
        sim.step(args.window_steps)
        sim.sync()
        w_el = time.perf_counter() - t0
        out["window"] = {"steps": f"20-{20 + args.window_steps}", "ms_per_step": round(w_el / args.window_steps * 1e3, 4),
                         "particle_steps_per_s": np_ * args.window_steps / w_el}
        print(f"# window steps 20-{20 + args.window_steps}: {w_el / args.window_steps * 1e3:.4f} ms/step",
              file=sys.stderr, flush=True)
    del sim
    if args.copy_probe:
        copy_probe(1 << 30)

    if args.cpu_steps > 0:
        threads = args.cpu_threads or os.cpu_count()
        cpu_steps = max(1, int(args.cpu_steps * 100_000_000 // np_))     # bounded sample: ~the same CPU time
        print(f"# cpu baseline: {cpu_steps} FP32 CPU steps on {threads} threads", file=sys.stderr, flush=True)
        v, el_cpu, used = cpu_baseline(n, mode, dt, arrs, cpu_steps, threads, shift)
        out["cpu_baseline"] = {"value": v, "unit": "particle-steps/s", "cores": used, "kind": "port",
                               "sample": f"{cpu_steps} full steps of the same workload ({np_} particles, "
                                         f"{n}^2) after 1 warm-up step, {el_cpu:.1f}s; oracle/pm_cpu32.c "
                                         f"(int64 deposit into per-thread lattices allocated once, FP32 FFT "
                                         f"convolution, FP32 reference stencil and push; OpenMP), "
                                         f"host {cpu_model()}"}
    else:
        out["cpu_baseline"] = None
    if args.serving == "on":
        out["serving"] = serving_bench(args.serving_requests, args.seed, args.serving_modes.split(","),
                                       args.serving_arrivals.split(","), args.serving_graphs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
