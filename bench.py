#!/usr/bin/env python
"""Benchmark: batched isotropic fast-sweeping eikonal solve on one MI355X.

Workload (the batched configuration): the 1001 x 4001 (dh 5 m) layered model,
256 surface shots 20 m apart, each solved on its own window of U{1200..4001}
columns (seed 3), FP32, one pass of four sweeps per shot (the reference's
schedule semantics), skewed-tile schedule replayed as one HIP graph.  A step
is one solve of all 256 shots, including initialisation and receiver picks,
with the model resident in HBM.

Reported beside it: the plain per-shot GPU solve (one launch per
anti-diagonal, the reference's schedule) on a sample of the same windows
(rate and spread), the same windows solved one shot at a time with the tiled
schedule, the drop-in per-shot driver loop (seisfat_eik_solve on the whole
model), the HBM roofline fraction of the sweep kernel, the CPU oracle timed on
samples of the same windows (one thread, and shot-parallel on the host cores),
traveltime errors of every config against the FP32 and FP64 oracle and t = r/v,
and two secondary workloads with their own rooflines: the batched anisotropic
(TTI qP) solve of 128 shots on 801 x 3001, and the adjoint-state misfit
gradient of 128 shots x 401 receivers on 301 x 1201.

Prints one JSON line.  `--gpus` above 1 is refused: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 12  # S read + T read + T write, 4 B each


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_sweep_pair_c3_256shots_traffic.txt")


def profile_figure(key):
    """A figure from the committed rocprofv3 summary of this workload: HBM
    bytes per sweep launch (FETCH_SIZE / WRITE_SIZE passes calibrated against
    a device copy) or the SQ pass's VALU-busy fraction."""
    try:
        kv = dict(l.split(":", 1) for l in open(TRAFFIC_FILE) if ":" in l)
        return float(kv[key])
    except (OSError, KeyError, ValueError):
        return None


def measured_traffic():
    return profile_figure("traffic_bytes_per_launch")


ANI_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "ani_c4_128shots_traffic.txt")
ADJ_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "adjoint_c5_128shots_traffic.txt")
MIG_COUNTERS_FILE = os.path.join(ROOT, "profiles", "kdm_migration_counters.txt")
VALU_PEAK_TIPS = 1024 * 2.4e9 / 4 / 1e12  # wave64 VALU instructions per second, 256 CUs x 4 SIMDs


def profile_figures(path):
    """key: number pairs of a committed counter summary ({} if absent)."""
    out = {}
    try:
        for line in open(path):
            if ":" in line:
                k, v = line.split(":", 1)
                try:
                    out[k.strip()] = float(v)
                except ValueError:
                    pass
    except OSError:
        pass
    return out


def _cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return os.cpu_count() or 0, model


def cpu_baseline(w, batch, budget_s):
    """The oracle (the C++ restatement of the reference's single-pass solve)
    on windows of the workload: (a) one thread, on the first windows until
    `budget_s` seconds; (b) shot-parallel over the host cores this process may
    use (at most 16, the box's share), on the next windows for about the same
    time.  Every window solved is also compared with the GPU traveltimes bit
    for bit."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    from seisfat2d_amd import models
    done, cu, t_cpu, mismatches = 0, 0, 0.0, 0
    while t_cpu < budget_s and done < len(w.shots):
        sx, sz, x0, wd, _, _ = w.shots[done]
        t0 = time.perf_counter()
        T = pyoracle.solve_windows(w.s, w.dx, w.dz, [sx], [sz], [x0], [wd], nthreads=1)[0]
        t_cpu += time.perf_counter() - t0
        if not np.array_equal(T, batch.T(done)):
            mismatches += 1
        cu += models.cell_updates(w.nz, [wd])
        done += 1
    single = cu / t_cpu
    nproc, model = _cpu_info()
    threads = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc))
    # about budget_s of wall time: the single-thread rate x threads predicts it
    k0, k1, cu_mt = done, done, 0
    target = single * threads * budget_s
    while k1 < len(w.shots) and cu_mt < target:
        cu_mt += models.cell_updates(w.nz, [w.shots[k1][3]])
        k1 += 1
    if k1 == k0:  # every window went to the single-thread leg: time them again
        k0, k1, cu_mt = 0, done, cu
    sel = w.shots[k0:k1]
    t0 = time.perf_counter()
    Ts = pyoracle.solve_windows(w.s, w.dx, w.dz, [x[0] for x in sel], [x[1] for x in sel], [x[2] for x in sel],
                                [x[3] for x in sel], nthreads=threads)
    t_mt = time.perf_counter() - t0
    for k, T in zip(range(k0, k1), Ts):
        if not np.array_equal(T, batch.T(k)):
            mismatches += 1
    checked = done + (k1 - k0 if k0 == done else 0)
    return {"value": cu_mt / t_mt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"windows {k0}-{k1 - 1} of the {len(w.shots)}, one pass each, oracle/oracle.cpp (g++ -O3, "
                      f"FP32), shot-parallel on {threads} threads, {t_mt:.1f} s; single-thread leg on windows "
                      f"0-{done - 1}, {t_cpu:.1f} s",
            "shots_per_s": (k1 - k0) / t_mt, "nproc": nproc, "cpu_model": model,
            "single_core": {"value": single, "unit": "cell-updates/s", "cores": 1, "shots_per_s": done / t_cpu},
            "gpu_windows_bit_exact": f"{checked - mismatches}/{checked}"}


def bench_ani(steps):
    """128 full-model TTI qP shots on 801 x 3001 (dh 5 m): iso solve, qP
    slowness, qP solve and picks per shot, one HIP graph per step."""
    import torch
    from seisfat2d_amd import ANI, SCHED_TILED, Eikonal, models
    w = models.c4()
    e = Eikonal(w.nz, w.nx, w.dz, w.dx)
    e.set_slowness(w.s)
    e.set_stiffness(w.stiffness)
    b = e.batch(w.shots, w.rx, w.rz, type=ANI, schedule=SCHED_TILED)
    b.run()
    b.status()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    b.status()
    # device time of the solve phase (both solves and the qP slowness pass,
    # HIP events around them inside the run), mean of 3 runs
    ph = []
    for _ in range(3):
        b.run()
        ph.append(b.sweep_ms())
    b.status()
    phase_s = sum(ph) / len(ph) / 1e3
    ncu = models.cell_updates(w.nz, [s[3] for s in w.shots])
    cu = 2 * ncu
    # algorithmic bytes: 12 B per cell-update of both solves, and the qP pass's
    # 24 B per padded node and shot (T 4-neighbour read once, 6 x u16 stiffness
    # (shared, counted per shot), S read and written)
    nodes = (w.nz + 6) * (w.nx + 6) * len(w.shots)
    alg = BYTES_PER_CELL_UPDATE * cu + 24 * nodes
    achieved = alg / phase_s / 1e9
    tr = profile_figures(ANI_TRAFFIC_FILE)
    out = {"workload": f"batched TTI qP FSM, {len(w.shots)} full-model shots on {w.nz}x{w.nx} dh=5 m "
                       "(eps, delta, tilt smooth seeded fields), two single-pass solves + qP slowness per shot",
           "ms_per_step": dt * 1e3, "shots_per_s": len(w.shots) / dt, "cell_updates_per_s": cu / dt,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS,
                        "traffic": tr.get("traffic_bytes_per_step"),
                        "traffic_source": os.path.relpath(ANI_TRAFFIC_FILE, ROOT),
                        "scope": "solve phase (4 k_sweep_pair launches + k_ani_table + k_quasi_slowness)",
                        "algorithmic_bytes_per_step": alg, "phase_ms": phase_s * 1e3}}
    b.close()
    e.close()
    return out


def bench_adjoint(steps):
    """Misfit gradient of 128 shots x 401 receivers on 301 x 1201 (dh 10 m):
    recorded forward (fp32 batch + fp64 trajectory) and the adjoint sweeps."""
    import torch
    from seisfat2d_amd import Eikonal, models
    w = models.c5()
    e = Eikonal(w.nz, w.nx, w.dz, w.dx)
    e.set_slowness(w.extra["s_true"])
    b = e.batch(w.shots, w.rx, w.rz)
    b.run()
    obs = torch.from_numpy(b.picks()).cuda()
    e.set_slowness(w.s)
    g = torch.empty((w.nx, w.nz), dtype=torch.float64, device="cuda")
    t_rec = t_vjp = 0.0
    for it in range(steps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run_record()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res = torch.as_tensor(_DevPicks(b.device_picks_ptr(), len(w.rx)), device="cuda") - obs
        b.vjp_device(res.data_ptr(), g.data_ptr())
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if it > 0:
            t_rec += t1 - t0
            t_vjp += t2 - t1
    t_rec, t_vjp = t_rec / steps, t_vjp / steps
    b.status()
    # algorithmic bytes: the forward solve's 12 B per cell-update and the
    # adjoint's 16 B (T, S and lambda read, lambda written)
    cu = models.cell_updates(w.nz, [s[3] for s in w.shots])
    alg = (BYTES_PER_CELL_UPDATE + 16) * cu
    achieved = alg / (t_rec + t_vjp) / 1e9
    tr = profile_figures(ADJ_TRAFFIC_FILE)
    out = {"workload": f"misfit gradient, {len(w.shots)} shots x 401 receivers on {w.nz}x{w.nx} dh=10 m "
                       "(v(z) start model, picks of v(z)+4 anomalies)",
           "ms_record": t_rec * 1e3, "ms_vjp": t_vjp * 1e3, "ms_gradient": (t_rec + t_vjp) * 1e3,
           "shots_per_s": len(w.shots) / (t_rec + t_vjp), "grad_norm": float(g.norm()),
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": tr.get("traffic_bytes_per_gradient"),
                        "traffic_source": os.path.relpath(ADJ_TRAFFIC_FILE, ROOT),
                        "scope": "whole gradient (recorded forward + vector-Jacobian product), wall clock",
                        "algorithmic_bytes_per_gradient": alg}}
    b.close()
    e.close()
    return out


def calibration_copy():
    """Counter calibration for profiler runs: a plain device copy of exactly
    256 MiB (reads 256 MiB, writes 256 MiB)."""
    import torch
    src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"profile: calibration copy of {src.numel() * 4} bytes done")


def drop_in_per_shot(nshots=10):
    """The reference's per-shot driver loop through the drop-in C ABI
    (seisfat_eik_solve, whose one-shot batch is built once and reused): shots
    on the whole 1001 x 4001 model, surface receivers every 20 m picked on the
    device; once without and once with the D2H copy of the padded traveltimes
    that Modeling::compute_seismogram makes (src/modeling/modeling.cu:165)."""
    import numpy as np
    from seisfat2d_amd import SCHED_TILED, Eikonal, models
    w = models.c2()
    e = Eikonal(w.nz, w.nx, w.dz, w.dx)
    e.set_slowness(w.s)
    xs = np.linspace(100.0, (w.nx - 1) * w.dx - 100.0, nshots)
    e.solve(float(xs[0]), 0.0, schedule=SCHED_TILED, rx=w.rx, rz=w.rz, return_T=False)  # builds the batch
    out = {}
    for with_T in (False, True):
        t0 = time.perf_counter()
        for x in xs:
            e.solve(float(x), 0.0, schedule=SCHED_TILED, rx=w.rx, rz=w.rz, return_T=with_T)
        dt = time.perf_counter() - t0
        out["shots_per_s_with_T_copy" if with_T else "shots_per_s"] = nshots / dt
    e.close()
    out["workload"] = (f"{nshots} shots on the whole {w.nz}x{w.nx} model, one seisfat_eik_solve per shot "
                       f"(tiled schedule), {len(w.rx)} receivers")
    return out


def bench_tables(ntables=64):
    """Traveltime tables for migration (Migration::set_src_travel_times,
    src/migration/migration.cu:116-134): padded grids of whole-model solves
    on the 1001 x 4001 model from `ntables` surface positions, as one batch;
    tables/s with the tables left in HBM, and with each copied to host memory
    (the reference writes every table to a file from host memory)."""
    import numpy as np
    import torch
    from seisfat2d_amd import SCHED_TILED, Eikonal, models
    w = models.c2()
    e = Eikonal(w.nz, w.nx, w.dz, w.dx)
    e.set_slowness(w.s)
    xs = np.linspace(0.0, (w.nx - 1) * w.dx, ntables)
    b = e.batch([(float(x), 0.0, 0, w.nx, 0, 0) for x in xs], schedule=SCHED_TILED)
    b.run()
    b.status()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        b.run()
    torch.cuda.synchronize()
    t_dev = (time.perf_counter() - t0) / reps
    b.status()
    host = np.empty((w.nx + 6, w.nz + 6), np.float32)
    t0 = time.perf_counter()
    b.run()
    for k in range(ntables):
        host[...] = b.T(k)
    t_host = time.perf_counter() - t0
    b.close()
    e.close()
    return {"workload": f"{ntables} traveltime tables (padded {w.nz + 6}x{w.nx + 6} grids, one per surface position) "
                        f"on the whole {w.nz}x{w.nx} model, one batch",
            "tables_per_s": ntables / t_dev, "tables_per_s_with_host_copy": ntables / t_host}


def bench_migration(steps=3, cpu_traces=200, least_squares=True):
    """Kirchhoff depth migration with the reference's sample migration
    parameters (models.migration: 1001 x 1001 at 10 m, 51 sources, 201
    receivers, offsets below 2000 m, nt 1001 at 1 ms): traveltime tables of
    every source and receiver position (one batch each, left in HBM), then
    per domain the adjoint operator (the migration's kernel, HIP events on
    torch's stream, median of `steps`), the whole migration (gather, adjoint
    convolution, adjoint operator; wall clock), the forward operator and
    (least_squares) three least-squares iterations (wall clock).
    Reports node-trace pairs per second (the reference launches one thread
    per padded node and trace) and the pairs that reach the trace, and the
    oracle timed on `cpu_traces` traces of the image-domain adjoint on one
    host core."""
    import numpy as np
    import torch
    from seisfat2d_amd import Eikonal, models
    from seisfat2d_amd import migration as mig
    w = models.migration()
    P = w.extra["params"]
    xs, zs, xr, zr = w.extra["xs"], w.extra["zs"], w.rx, w.rz
    e = Eikonal(w.nz, w.nx, w.dz, w.dx)
    e.set_slowness(w.s)
    t0 = time.perf_counter()
    ts = e.traveltime_tables(xs, zs, chunk=len(xs))
    tr = e.traveltime_tables(xr, zr, chunk=len(xr))
    t_tables = time.perf_counter() - t0
    out = {"workload": (f"Kirchhoff depth migration, {w.nz}x{w.nx} dh=10 m layered model, {len(xs)} sources every "
                        f"200 m, {len(xr)} receivers every 50 m, offsets < {P['max_offset']:.0f} m, nt {P['nt']} "
                        f"at {P['dt'] * 1e3:.0f} ms, fmax {P['fmax']:.0f} Hz, aperture {P['aperture']:.0f} deg, "
                        f"angles to {P['max_angle']:.0f} deg every {P['da']:.0f} deg"),
           "tables_s_with_host_copies": t_tables}

    def events_ms(fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ms = []
        for _ in range(steps):
            a.record()
            fn()
            b.record()
            b.synchronize()
            ms.append(a.elapsed_time(b))
        return float(np.median(ms))

    g = torch.Generator(device="cuda").manual_seed(0)
    for domain, name in ((mig.IMAGE, "image"), (mig.ANGLE, "angle")):
        K = mig.Kirchhoff(e, xs, zs, xr, zr, P["nt"], P["dt"], P["aperture"], P["max_offset"], P["fmax"],
                          domain=domain, da=P["da"], max_angle=P["max_angle"], tables=(ts, tr))
        seismic = torch.randn((K.nsrc, K.nrec, K.nt), device="cuda", generator=g)
        if domain == mig.IMAGE:
            # the table quantities and chunk minima of all the tables (seisfat_kdm_set_tables)
            ts_d, tr_d = K.set_tables(ts, tr)
            out["set_tables_ms"] = events_ms(lambda: K.set_tables(ts_d, tr_d))
            del ts_d, tr_d
        d = K.convolve(K.gather(seismic), adjoint=True)
        model = K.adjoint(d)
        torch.cuda.synchronize()
        adj_ms = events_ms(lambda: K.adjoint(d, model.zero_()))
        fwd_ms = events_ms(lambda: K.forward(model, d.zero_()))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            K.migrate(seismic)
        torch.cuda.synchronize()
        mig_ms = (time.perf_counter() - t0) / steps * 1e3
        # least-squares migration (IDLSKDM / ADLSKDM): three iterations, each two
        # forwards and an adjoint with their convolutions and the step's sums
        lsm_ms = None
        if least_squares:
            smooth = 1e-3 if domain == mig.ANGLE else 0.0
            K.lsm(seismic, max_it=3, smooth=smooth)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            K.lsm(seismic, max_it=3, smooth=smooth)
            torch.cuda.synchronize()
            lsm_ms = (time.perf_counter() - t0) * 1e3
        pairs = K.ntr * w.nz * w.nx
        r = {"traces": K.ntr, "model_shape": list(K.model_shape), "adjoint_ms": adj_ms, "forward_ms": fwd_ms,
             "migration_ms": mig_ms, "traces_per_s": K.ntr / (mig_ms / 1e3),
             "least_squares_3_iterations_ms": lsm_ms,
             "adjoint_pairs_per_s": pairs / (adj_ms / 1e3), "forward_pairs_per_s": pairs / (fwd_ms / 1e3)}
        if domain == mig.IMAGE:
            # node-trace pairs that reach the trace (time index < nt), counted on the device
            dt32 = torch.tensor(np.float32(P["dt"]), device="cuda")
            Ts = torch.from_numpy(ts[:, 3:-3, 3:-3].copy()).cuda()
            Tr = torch.from_numpy(tr[:, 3:-3, 3:-3].copy()).cuda()
            reach = 0
            for t in range(K.ntr):
                reach += int((torch.floor((Ts[K.src[t]] + Tr[K.rec[t]]) / dt32) < P["nt"]).sum())
            del Ts, Tr
            r["pairs_reaching"] = reach
            r["reaching_fraction"] = reach / pairs
            r["adjoint_reaching_pairs_per_s"] = reach / (adj_ms / 1e3)
            out["cpu_baseline"] = _migration_cpu(K, ts, tr, w, P, cpu_traces, reach / pairs)
        out[name] = r
        K.close()
    # the same survey with an 8 s record (nt 8001): most node-trace pairs
    # reach the trace, so the adjoint is bound by the weights' evaluation
    # and the table quantities' traffic, not by the passes over chunks
    nt_deep = 8001
    K = mig.Kirchhoff(e, xs, zs, xr, zr, nt_deep, P["dt"], P["aperture"], P["max_offset"], P["fmax"],
                      domain=mig.IMAGE, tables=(ts, tr))
    d = torch.randn((K.ntr, K.nt), device="cuda", generator=g)
    model = K.adjoint(d)
    torch.cuda.synchronize()
    adj_ms = events_ms(lambda: K.adjoint(d, model.zero_()))
    fwd_ms = events_ms(lambda: K.forward(model, d.zero_()))
    dt32 = torch.tensor(np.float32(P["dt"]), device="cuda")
    Ts = torch.from_numpy(ts[:, 3:-3, 3:-3].copy()).cuda()
    Tr = torch.from_numpy(tr[:, 3:-3, 3:-3].copy()).cuda()
    reach = 0
    for t in range(K.ntr):
        reach += int((torch.floor((Ts[K.src[t]] + Tr[K.rec[t]]) / dt32) < nt_deep).sum())
    del Ts, Tr
    pairs = K.ntr * w.nz * w.nx
    # the launch's busiest resource is VALU issue (counters; its fetch past L2
    # comes next): its wave-level VALU instructions (the committed profile's
    # per-pair count times the pairs that reach) over the launch time, against
    # the issue rate of 1024 SIMDs at one wave64 instruction per 4 cycles at
    # 2.4 GHz
    prof = profile_figures(MIG_COUNTERS_FILE)
    per_pair = prof.get("valu_instructions_per_pair")
    valu = per_pair * reach / 64 if per_pair else None
    achieved = valu / (adj_ms / 1e3) / 1e12 if valu else None
    fetch = prof.get("fetch_bytes_per_launch")
    out["image_deep_record"] = {
        "nt": nt_deep, "traces": K.ntr, "adjoint_ms": adj_ms, "forward_ms": fwd_ms, "pairs_reaching": reach,
        "reaching_fraction": reach / pairs, "adjoint_pairs_per_s": pairs / (adj_ms / 1e3),
        "adjoint_reaching_pairs_per_s": reach / (adj_ms / 1e3),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TIPS, "unit": "T wave-instr/s",
                     "frac": achieved / VALU_PEAK_TIPS if achieved else None,
                     "traffic": fetch, "hbm_GBps_at_traffic": fetch / (adj_ms / 1e3) / 1e9 if fetch else None,
                     "valu_instructions_per_pair": per_pair,
                     "scope": "k_kdm_adjoint<image> launch, HIP events on torch's stream",
                     "source": os.path.relpath(MIG_COUNTERS_FILE, ROOT)}}
    K.close()
    e.close()
    return out


def _migration_cpu(K, ts, tr, w, P, ntr, reach_frac):
    """The oracle (the reference's per-trace kernel restated on the CPU) on
    the first `ntr` traces of the image-domain adjoint: one thread, and the
    host cores this process may use (at most 16) over image columns."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as o
    S = o.expand_boundary(w.s)
    d = np.random.default_rng(0).standard_normal((ntr, K.nt)).astype(np.float32)
    threads = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1))
    rates = {}
    for n in sorted({1, threads}):
        m = np.zeros(K.model_shape, np.float32)
        t0 = time.perf_counter()
        o.kdm(0, True, S, ts, tr, K.src[:ntr], K.rec[:ntr], K.cmp[:ntr], None, d, m, P["dt"], w.dx, w.dz,
              P["aperture"], 1.0, 1, nthreads=n)
        rates[n] = ntr / (time.perf_counter() - t0)
    return {"value": rates[threads] * w.nz * w.nx, "unit": "node-trace pairs/s", "cores": threads, "kind": "port",
            "sample": f"first {ntr} traces of the image-domain adjoint, {threads} threads over image columns",
            "traces_per_s": rates[threads],
            "single_core": {"value": rates[1] * w.nz * w.nx, "traces_per_s": rates[1]}}


def traveltime_errors():
    """Max relative traveltime error over interior nodes (T_ref > 0) of the GPU
    solve against the FP32 oracle (the reference's CPU path restated; the
    parity gate, 1e-5) and the FP64 oracle, per config, and of C1's
    constant-velocity variant against t = r/v."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as o
    from seisfat2d_amd import ANI, Eikonal, models

    def rel(a, ref):
        a = a[3:-3, 3:-3].astype(np.float64)
        ref = ref[3:-3, 3:-3].astype(np.float64)
        m = ref > 0
        return float(np.max(np.abs(a[m] - ref[m]) / ref[m]))

    out = {}
    for name, w, (sx, sz) in (("C1", models.c1(), (5000.0, 0.0)), ("C2", models.c2(), (10000.0, 0.0)),
                              ("C5", models.c5(), (models.c5().shots[37][0], 0.0))):
        e = Eikonal(w.nz, w.nx, w.dz, w.dx)
        e.set_slowness(w.s)
        T, _ = e.solve(sx, sz)
        e.close()
        S32 = o.expand_boundary(w.s)
        r = {"vs_fp32_oracle": rel(T, o.solve_iso(S32, w.dx, w.dz, sx, sz)),
             "vs_fp64_oracle": rel(T, o.solve_iso(o.expand_boundary(w.s.astype(np.float64)), w.dx, w.dz, sx, sz))}
        if name == "C1":
            wc = models.c1(constant=True)
            e = Eikonal(wc.nz, wc.nx, wc.dz, wc.dx)
            e.set_slowness(wc.s)
            Tc, _ = e.solve(sx, sz)
            e.close()
            x = (np.arange(wc.nx + 6) - 3) * wc.dx
            z = (np.arange(wc.nz + 6) - 3) * wc.dz
            r["constant_velocity_vs_r_over_v"] = rel(Tc, np.hypot(x[:, None] - sx, z[None, :] - sz) / 2000.0)
        out[name] = r
    # C3: two windows of the batch workload against FP32 and FP64
    w = models.c3()
    e = Eikonal(w.nz, w.nx, w.dz, w.dx)
    e.set_slowness(w.s)
    worst32 = worst64 = 0.0
    for k in (0, 200):
        sx, sz, x0, wd, _, _ = w.shots[k]
        b = e.batch([(sx, sz, x0, wd, 0, 0)])
        b.run()
        T = b.T(0)
        b.close()
        crop = np.ascontiguousarray(w.s[x0:x0 + wd])
        worst32 = max(worst32, rel(T, o.solve_iso(o.expand_boundary(crop), w.dx, w.dz, sx - x0 * w.dx, sz)))
        worst64 = max(worst64, rel(T, o.solve_iso(o.expand_boundary(crop.astype(np.float64)), w.dx, w.dz,
                                                  sx - x0 * w.dx, sz)))
    e.close()
    out["C3"] = {"vs_fp32_oracle": worst32, "vs_fp64_oracle": worst64, "sample": "windows 0 and 200"}
    # C4: one shot of the TTI batch against the FP32 oracle (the oracle's
    # anisotropic path is FP32 only)
    w = models.c4(nshots=2)
    e = Eikonal(w.nz, w.nx, w.dz, w.dx)
    e.set_slowness(w.s)
    e.set_stiffness(w.stiffness)
    sx = w.shots[1][0]
    T, _ = e.solve(sx, 0.0, type=ANI)
    e.close()
    Cq, mn, mx = [], [], []
    for c in w.stiffness:
        u, a, bb = o.compress(o.expand_boundary(c))
        Cq.append(u)
        mn.append(a)
        mx.append(bb)
    out["C4"] = {"vs_fp32_oracle": rel(T, o.solve_ani(o.expand_boundary(w.s), w.dx, w.dz, sx, 0.0, Cq, mn, mx)[0]),
                 "sample": "shot 2 of 2 evenly spaced"}
    return out


class _DevPicks:
    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<f4", "data": (int(ptr), False),
                                         "version": 2}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--shots", type=int, default=256)
    ap.add_argument("--plain-shots", type=int, default=8, help="windows timed with the per-diagonal schedule")
    ap.add_argument("--errors", type=int, default=1, help="report traveltime errors of every config against "
                                                          "the FP32 and FP64 oracle and against t = r/v")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU oracle budget (0 = skip)")
    ap.add_argument("--extras", type=int, default=1, help="also time the anisotropic and adjoint workloads")
    ap.add_argument("--tune", type=int, default=1, help="tune the batch's schedule options on this device first "
                                                        "(seisfat2d_amd.tune, not cached)")
    ap.add_argument("--opts", type=str, default=None,
                    help='schedule options as JSON (e.g. \'{"wpcu": 8}\'); no tuning')
    ap.add_argument("--profile", action="store_true",
                    help="profiler run: the batched steps only, plus a 256 MiB device copy that calibrates "
                         "per-dispatch byte counters")
    ap.add_argument("--only", choices=["ani", "adjoint", "migration"],
                    help="profiler run of one secondary workload: its steps only, printed as one JSON line")
    args = ap.parse_args()
    if args.profile:
        args.plain_shots, args.cpu_seconds = 0, 0.0
        args.tune = 0  # a profile holds the timed kernels only
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(shots would shard across GPUs with no exchange; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from seisfat2d_amd import SCHED_PLAIN, SCHED_TILED, Eikonal, models

    torch.cuda.init()
    if args.only:
        run = {"ani": bench_ani, "adjoint": bench_adjoint,
               "migration": lambda k: bench_migration(k, least_squares=not args.profile)}[args.only]
        print(json.dumps(run(max(1, args.steps))))
        if args.profile:
            calibration_copy()
        return
    dev = torch.cuda.get_device_properties(0)
    log(f"device {dev.name} CUs {dev.multi_processor_count}")
    w = models.c3(nshots=args.shots)
    widths = [s[3] for s in w.shots]
    cu = models.cell_updates(w.nz, widths)
    e = Eikonal(w.nz, w.nx, w.dz, w.dx)
    e.set_slowness(w.s)
    tuned = None
    if args.opts:
        opts = json.loads(args.opts)
    elif args.tune:
        from seisfat2d_amd import tune
        tuned = tune.tune(e, w.shots, w.rx, w.rz, force=True, cache=False)
        opts = tuned["opts"]
        log(f"tuned schedule {opts} ({len(tuned['trials'])} candidates, best sweep phase {tuned['ms']:.2f} ms)")
    else:
        opts = None
    b = e.batch(w.shots, w.rx, w.rz, schedule=SCHED_TILED, opts=opts)
    sched = b.opts
    log(f"batch: {len(w.shots)} shots, {b.num_launches} launches/step ({b.num_sweep_launches} sweep), "
        f"schedule {sched}")

    for _ in range(args.warmup):
        b.run()
    b.status()  # raises if a persistent sweep flagged a failure
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    b.status()  # the timed runs completed without a device-side failure
    tiles_total, tiles_skipped = b.tile_counts()
    ms_step = dt / args.steps * 1e3
    value = cu * args.steps / dt
    shots_per_s = len(w.shots) * args.steps / dt
    log(f"batched: {ms_step:.2f} ms/step, {shots_per_s:.0f} shots/s, {value / 1e9:.1f} G cell-updates/s")

    # sweep-phase device time from the HIP events inside the graph (average of 5 runs)
    sweep = []
    for _ in range(5):
        b.run()
        sweep.append(b.sweep_ms())
    sweep_ms = float(np.mean(sweep))
    nlaunch = b.num_sweep_launches
    # two launches: the fused sweep pairs (shots spread over the XCDs); four:
    # one persistent launch per sweep
    sweep_kernel = "k_sweep_pair" if nlaunch == 2 else "k_sweep_persist"
    bytes_per_launch = BYTES_PER_CELL_UPDATE * cu / nlaunch
    avg_launch_s = sweep_ms / 1e3 / nlaunch
    achieved = bytes_per_launch / avg_launch_s / 1e9
    log(f"sweep phase {sweep_ms:.2f} ms over {nlaunch} launches -> {achieved:.0f} GB/s algorithmic")

    if args.profile:
        calibration_copy()
        b.close()
        e.close()
        return

    # per-shot comparisons on the first windows: the reference's schedule
    # (one launch per anti-diagonal) and the tiled one, one window at a time
    nps = max(1, min(args.plain_shots, len(w.shots)))

    def per_shot(sched):
        rates = []
        for k in range(nps):
            bp = e.batch([w.shots[k][:4] + (0, 0)], schedule=sched)
            bp.run()
            bp.status()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            bp.run()
            torch.cuda.synchronize()
            rates.append(1.0 / (time.perf_counter() - t1))
            bp.status()
            bp.close()
        rates = np.array(rates)
        return nps / np.sum(1.0 / rates), {"min": float(rates.min()), "median": float(np.median(rates)),
                                           "max": float(rates.max()), "windows": nps}

    plain_shots_per_s, plain_spread = per_shot(SCHED_PLAIN)
    single_shots_per_s, single_spread = per_shot(SCHED_TILED)
    log(f"per-shot: plain {plain_shots_per_s:.1f} shots/s ({plain_spread}), tiled {single_shots_per_s:.1f} shots/s")
    drop_in = drop_in_per_shot()
    tables = bench_tables()
    log(f"tables: {tables['tables_per_s']:.0f}/s in HBM, {tables['tables_per_s_with_host_copy']:.0f}/s copied to host")
    log(f"drop-in per-shot driver (whole C2 model): {drop_in['shots_per_s']:.1f} shots/s, "
        f"{drop_in['shots_per_s_with_T_copy']:.1f} with the T copy")

    extras = {}
    if args.extras:
        extras["ani_c4"] = bench_ani(max(1, min(args.steps, 5)))
        log(f"ani: {extras['ani_c4']['ms_per_step']:.1f} ms/step, {extras['ani_c4']['shots_per_s']:.0f} shots/s")
        extras["adjoint_c5"] = bench_adjoint(max(1, min(args.steps, 5)))
        a = extras["adjoint_c5"]
        log(f"adjoint: record {a['ms_record']:.1f} ms + vjp {a['ms_vjp']:.1f} ms per gradient of 128 shots")
        extras["migration"] = bench_migration(max(1, min(args.steps, 5)))
        m = extras["migration"]
        log(f"migration: image adjoint {m['image']['adjoint_ms']:.2f} ms, migration {m['image']['migration_ms']:.1f} "
            f"ms for {m['image']['traces']} traces; angle adjoint {m['angle']['adjoint_ms']:.2f} ms; 3 least-squares "
            f"iterations {m['image']['least_squares_3_iterations_ms']:.1f} ms (image), "
            f"{m['angle']['least_squares_3_iterations_ms']:.1f} ms (angle)")

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(w, b, args.cpu_seconds)
        log(f"cpu oracle: {cpu['value'] / 1e6:.1f} M cell-updates/s on {cpu['cores']} threads "
            f"({cpu['single_core']['value'] / 1e6:.1f} M on one), windows bit-exact {cpu['gpu_windows_bit_exact']}")
    errors = traveltime_errors() if args.errors else None
    if errors:
        log(f"errors: {errors}")

    out = {
        "metric": "eikonal cell-updates/s & shots/s, batched vs per-shot, at ≤1e-5 rel err; HBM %",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"batched iso FSM, {len(w.shots)} shots on windows 1200-4001 wide of the "
                               f"1001x4001 dh=5 m layered model (seeded), one pass of 4 sweeps, tiled "
                               f"schedule ({sweep_kernel} launches) on 64x64 tiles with schedule options "
                               f"{json.dumps(sched, sort_keys=True)}"
                               f"{' (tuned on this device)' if tuned else ''}, one HIP graph per step",
                   "nz": w.nz, "nx": w.nx, "shots": len(w.shots), "cell_updates_per_step": cu,
                   "schedule": sched},
        "schedule_tuning": ({"trials": tuned["trials"], "best_sweep_ms": tuned["ms"]} if tuned else None),
        "tiles_per_step": tiles_total,
        "tiles_skipped_per_step": tiles_skipped,
        "shots_per_s": shots_per_s,
        "per_shot_plain_shots_per_s": plain_shots_per_s,
        "per_shot_plain_spread": plain_spread,
        "per_shot_tiled_shots_per_s": single_shots_per_s,
        "per_shot_tiled_spread": single_spread,
        "batched_vs_plain_per_shot": shots_per_s / plain_shots_per_s,
        "batched_vs_plain_per_shot_worst": shots_per_s / plain_spread["max"],
        "drop_in_per_shot": drop_in,
        "traveltime_tables": tables,
        "rel_err_vs_oracle": 0.0 if (cpu and cpu["gpu_windows_bit_exact"].split("/")[0] ==
                                    cpu["gpu_windows_bit_exact"].split("/")[1]) else None,
        "errors": errors,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     # the same over the tiles the kernels ran (tiles with only
                     # not-reached inputs are not run where exact)
                     "frac_run_tiles": achieved / HBM_PEAK_GBS * (1.0 - tiles_skipped / max(tiles_total, 1)),
                     "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT),
                     # the binding resource: SQ_ACTIVE_INST_VALU over the sweep kernels
                     "valu_busy": profile_figure("valu_busy"),
                     "kernel": sweep_kernel, "launches": nlaunch,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "avg_launch_us": avg_launch_s * 1e6},
        "cpu_baseline": cpu,
        **extras,
    }
    print(json.dumps(out))
    b.close()
    e.close()


if __name__ == "__main__":
    main()
