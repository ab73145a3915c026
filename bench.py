#!/usr/bin/env python
"""Benchmark: batched isotropic fast-sweeping eikonal solve on one MI355X.

Workload (the batched configuration): the 1001 x 4001 (dh 5 m) layered model,
256 surface shots 20 m apart, each solved on its own window of U{1200..4001}
columns (seed 3), FP32, one pass of four sweeps per shot (the reference's
schedule semantics), skewed-tile schedule replayed as one HIP graph.  A step
is one solve of all 256 shots, including initialisation and receiver picks,
with the model resident in HBM.

Reported beside it: the plain per-shot GPU solve (one launch per
anti-diagonal, the reference's schedule) on a sample of the same windows, the
same windows solved one shot at a time with the tiled schedule, the HBM
roofline fraction of the sweep kernel, the CPU oracle timed on a sample of
the same windows, and two secondary workloads: the batched anisotropic (TTI
qP) solve of 128 shots on 801 x 3001, and the adjoint-state misfit gradient
of 128 shots x 401 receivers on 301 x 1201.

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


def cpu_baseline(w, batch, budget_s):
    """Oracle (single-threaded C++ restatement) on the first windows of the
    workload until `budget_s` seconds of CPU work; also checks those windows
    against the GPU traveltimes bit for bit."""
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
    return {"value": cu / t_cpu, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"first {done} of the {len(w.shots)} windows, one pass each, oracle/oracle.cpp "
                      f"(g++ -O3, FP32, single thread), {t_cpu:.1f} s",
            "shots_per_s": done / t_cpu, "gpu_windows_bit_exact": f"{done - mismatches}/{done}"}


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
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    b.status()
    cu = 2 * models.cell_updates(w.nz, [s[3] for s in w.shots])
    out = {"workload": f"batched TTI qP FSM, {len(w.shots)} full-model shots on {w.nz}x{w.nx} dh=5 m "
                       "(eps, delta, tilt smooth seeded fields), two single-pass solves + qP slowness per shot",
           "ms_per_step": dt * 1e3, "shots_per_s": len(w.shots) / dt, "cell_updates_per_s": cu / dt}
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
    out = {"workload": f"misfit gradient, {len(w.shots)} shots x 401 receivers on {w.nz}x{w.nx} dh=10 m "
                       "(v(z) start model, picks of v(z)+4 anomalies)",
           "ms_record": t_rec * 1e3, "ms_vjp": t_vjp * 1e3, "ms_gradient": (t_rec + t_vjp) * 1e3,
           "shots_per_s": len(w.shots) / (t_rec + t_vjp), "grad_norm": float(g.norm())}
    b.close()
    e.close()
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
    ap.add_argument("--plain-shots", type=int, default=3, help="windows timed with the per-diagonal schedule")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU oracle budget (0 = skip)")
    ap.add_argument("--extras", type=int, default=1, help="also time the anisotropic and adjoint workloads")
    ap.add_argument("--profile", action="store_true",
                    help="profiler run: the batched steps only, plus a 256 MiB device copy that calibrates "
                         "per-dispatch byte counters")
    ap.add_argument("--only", choices=["ani", "adjoint"],
                    help="profiler run of one secondary workload: its steps only, printed as one JSON line")
    args = ap.parse_args()
    if args.profile:
        args.plain_shots, args.cpu_seconds = 0, 0.0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this build has no multi-GPU path "
              "(shots would shard across GPUs with no exchange; see DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from seisfat2d_amd import SCHED_PLAIN, SCHED_TILED, Eikonal, models

    torch.cuda.init()
    if args.only:
        run = bench_ani if args.only == "ani" else bench_adjoint
        print(json.dumps(run(max(1, args.steps))))
        return
    dev = torch.cuda.get_device_properties(0)
    log(f"device {dev.name} CUs {dev.multi_processor_count}")
    w = models.c3(nshots=args.shots)
    widths = [s[3] for s in w.shots]
    cu = models.cell_updates(w.nz, widths)
    e = Eikonal(w.nz, w.nx, w.dz, w.dx)
    e.set_slowness(w.s)
    b = e.batch(w.shots, w.rx, w.rz, schedule=SCHED_TILED)
    log(f"batch: {len(w.shots)} shots, {b.num_launches} launches/step ({b.num_sweep_launches} sweep)")

    for _ in range(args.warmup):
        b.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
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
        # counter calibration: a plain copy of exactly 256 MiB (reads 256 MiB, writes 256 MiB)
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"profile: calibration copy of {src.numel() * 4} bytes done")
        b.close()
        e.close()
        return

    # per-shot comparisons on the first windows
    nps = max(1, min(args.plain_shots, len(w.shots)))
    t_plain = 0.0
    for k in range(nps):
        bp = e.batch([w.shots[k][:4] + (0, 0)], schedule=SCHED_PLAIN)
        bp.run()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        bp.run()
        torch.cuda.synchronize()
        t_plain += time.perf_counter() - t1
        bp.close()
    plain_shots_per_s = nps / t_plain
    t_single = 0.0
    for k in range(nps):
        bs = e.batch([w.shots[k][:4] + (0, 0)], schedule=SCHED_TILED)
        bs.run()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        bs.run()
        torch.cuda.synchronize()
        t_single += time.perf_counter() - t1
        bs.close()
    single_shots_per_s = nps / t_single
    log(f"per-shot: plain {plain_shots_per_s:.1f} shots/s, tiled {single_shots_per_s:.1f} shots/s")

    extras = {}
    if args.extras:
        extras["ani_c4"] = bench_ani(max(1, min(args.steps, 5)))
        log(f"ani: {extras['ani_c4']['ms_per_step']:.1f} ms/step, {extras['ani_c4']['shots_per_s']:.0f} shots/s")
        extras["adjoint_c5"] = bench_adjoint(max(1, min(args.steps, 5)))
        a = extras["adjoint_c5"]
        log(f"adjoint: record {a['ms_record']:.1f} ms + vjp {a['ms_vjp']:.1f} ms per gradient of 128 shots")

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(w, b, args.cpu_seconds)
        log(f"cpu oracle: {cpu['value'] / 1e6:.1f} M cell-updates/s, windows bit-exact {cpu['gpu_windows_bit_exact']}")

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
                               f"schedule ({sweep_kernel} launches), one HIP graph per step",
                   "nz": w.nz, "nx": w.nx, "shots": len(w.shots), "cell_updates_per_step": cu},
        "shots_per_s": shots_per_s,
        "per_shot_plain_shots_per_s": plain_shots_per_s,
        "per_shot_tiled_shots_per_s": single_shots_per_s,
        "batched_vs_plain_per_shot": shots_per_s / plain_shots_per_s,
        "rel_err_vs_oracle": 0.0 if (cpu and cpu["gpu_windows_bit_exact"].split("/")[0] ==
                                    cpu["gpu_windows_bit_exact"].split("/")[1]) else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
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
