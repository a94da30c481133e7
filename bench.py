"""Benchmark of the fused moment-representation step on one MI355X.

Workload (default): D3Q19 Taylor-Green vortex, 256^3, triply periodic,
FP32 moments, u0 = 0.05, nu = 0.01 (tau = 0.53), 8x8x8 tiles. One "step" is
one lattice update of all 16.8 M nodes (one kernel launch).

Prints one JSON line: whole-job MLUPS (wall clock around the synchronized
timed region), the roofline object of the step kernel (algorithmic bytes per
launch / HIP-event-timed mean launch duration on the lattice's stream) and a
CPU baseline (the FP64 oracle on a bounded sample, host cores).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mr-lbm_amd"))

METRIC = "MLUPS & %HBM roofline, D3Q19 256³ fp32 vs fp16 moments; p50/p99 solve latency"
PEAK_HBM_GBS = 8000.0
# best published MLUPS of this path (D3Q19 FP32 256^3): A100, README.md:141
REF_PUBLISHED_MLUPS = 11390.0


def traffic_from_profiles(workload_key):
    """HBM bytes per launch measured with rocprofv3 PMC counters, if a
    summary for this workload is committed under profiles/."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload_key}.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def cpu_baseline(target_s, n=256):
    """FP64 oracle on the host cores, on the benchmark workload itself (TGV
    n^3, D3Q19, tau 0.53): OpenMP over all threads for about target_s
    seconds, then one thread for about a third of that (the serial figure)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle

    L = _oracle.lib()
    threads = L.orc_get_threads()
    o = _oracle.Oracle(n, n, n, q=19, tau=0.53)
    o.init_taylor_green(0.05)
    t1 = o.step_timed(1)  # calibration (also absorbs the thread pool's start-up)
    steps = max(2, min(5000, int(target_s / max(t1, 1e-6))))
    t = o.step_timed(steps)
    L.orc_set_threads(1)
    try:
        t1s = o.step_timed(1)
        steps_s = max(2, int(target_s / 3 / max(t1s, 1e-6)))
        ts = o.step_timed(steps_s)
    finally:
        L.orc_set_threads(threads)
    nodes = n ** 3
    return {
        "value": nodes * steps / t / 1e6,
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"FP64 oracle, D3Q19 TGV {n}^3, {steps} steps ({t:.1f} s), "
                  f"OpenMP {threads} threads",
        "serial_mlups": nodes * steps_s / ts / 1e6,
        "serial_sample": f"same lattice, {steps_s} steps ({ts:.1f} s), 1 thread",
    }


class SplitMix64:
    """Seeded generator for the synthetic workloads (own implementation so the
    draws do not depend on a library's distribution code)."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, a, b):
        return a + (b - a) * (self.next() >> 11) / float(1 << 53)

    def choice(self, seq):
        return seq[self.next() % len(seq)]


def batching_cavities(n_lat=64, steps=2000, seed=2024, check=4):
    """64 lid-driven cavities (edges from {32,48,64,96,128}, Re ~ U[50,400],
    lid 0.05, tau = 0.5 + 3 u N / Re) in one batch with hipGraph-captured
    steps and an autotuned tile, against the same lattices stepped one after
    another with one launch per step."""
    import numpy as np
    import mrlbm
    from mrlbm import geometry

    rng = SplitMix64(seed)
    lid = 0.05
    spec = []
    for _ in range(n_lat):
        n = rng.choice([32, 48, 64, 96, 128])
        re = rng.uniform(50.0, 400.0)
        spec.append((n, 0.5 + 3.0 * lid * n / re))
    dims = [(n, n, n) for n, _ in spec]
    shapes = ((8, 8, 8), (16, 8, 4), (32, 4, 4), (64, 4, 2))
    tile, tune_ms = mrlbm.autotune(dims, shapes=shapes, steps=50, walls=[(1, 1, 1)] * len(dims))
    types = {n: geometry.cavity(n) for n in set(n for n, _ in spec)}

    def make(n, tau):
        lat = mrlbm.Lattice((n, n, n), q=19, tau=tau, walls=(1, 1, 1), tile=tile, kernel=mrlbm.KERNEL_TILE)
        lat.set_bc_velocity(0, (lid, 0, 0))
        lat.set_node_types(types[n])
        lat.init_uniform()
        return lat

    nodes = sum(n ** 3 for n, _ in spec)
    gsteps = 50 if steps % 50 == 0 else 1
    batch = mrlbm.Batch(q=19, tile=tile)
    lats = [make(n, tau) for n, tau in spec]
    for lat in lats:
        batch.add(lat)
    batch.step(gsteps, graph_steps=gsteps)  # capture + warm-up
    t0 = time.perf_counter()
    ms_b = batch.time_steps(steps, graph_steps=gsteps)
    batch.synchronize()
    wall_b = time.perf_counter() - t0
    batched = [lats[i].moments() for i in range(check)]
    batch.close()
    for lat in lats:
        lat.close()

    ms_seq = 0.0
    identical = True
    t0 = time.perf_counter()
    for i, (n, tau) in enumerate(spec):
        lat = make(n, tau)
        lat.step(gsteps)
        ms_seq += lat.time_steps(steps)
        if i < check:
            identical &= bool(np.array_equal(lat.moments(), batched[i]))
        lat.close()
    wall_seq = time.perf_counter() - t0
    return {
        "workload": f"{n_lat} D3Q19 lid-driven cavities, edges from {{32,48,64,96,128}}, Re~U[50,400], "
                    f"lid 0.05, {steps} steps each (+{gsteps} warm-up), seed {seed}",
        "total_nodes": nodes,
        "tile": list(tile),
        "autotune_ms_50_steps": {"x".join(map(str, k)): v for k, v in tune_ms.items()},
        "batched_mlups": round(nodes * steps / (ms_b * 1e-3) / 1e6, 1),
        "sequential_mlups": round(nodes * steps / (ms_seq * 1e-3) / 1e6, 1),
        "speedup": round(ms_seq / ms_b, 3),
        "batched_ms": round(ms_b, 2),
        "sequential_ms": round(ms_seq, 2),
        "graph_steps": gsteps,
        "bit_identical_checked": check,
        "bit_identical": identical,
    }


def serving_ducts(n_req=2000, rate=20.0, seed=5, max_batch=6):
    """The duct request stream (2000 requests, Poisson rate 20/s, FP16
    storage, 200 GB budget) served in real time with continuous batching.
    At most max_batch lattices share the GPU: the admitted ducts saturate it
    already, and the cap serves them closer to arrival order than equal
    sharing would."""
    from mrlbm import serve

    reqs = serve.duct_requests(n_req, rate=rate, seed=seed)
    offered = sum(r.nx * r.ny * r.nz * r.steps for r in reqs) / reqs[-1].arrival_s / 1e6
    st = serve.run(reqs, storage=1, budget_gb=200.0, quantum=100, graph_steps=50, max_batch=max_batch, realtime=True)
    drift = max(abs(r.mass1 - r.mass0) for r in reqs if r.status == 0)
    return {
        "workload": f"{n_req} duct requests (seeded stream, seed {seed}), Poisson {rate}/s, "
                    f"FP16 moments, 200 GB budget, at most {max_batch} lattices in the batch, rounds of 100 steps "
                    "(graphs of 50)",
        "completed": sum(1 for r in reqs if r.status == 0),
        "failed": int(st["failed"]),
        "p50_latency_s": round(st["p50_s"], 3),
        "p99_latency_s": round(st["p99_s"], 3),
        "max_latency_s": round(st["max_s"], 3),
        "aggregate_mlups": round(st["mlups"], 1),
        "busy_mlups": round(st["mlups_busy"], 1),
        "utilisation": round(st["busy_s"] / st["wall_s"], 3),
        "offered_mlups": round(offered, 1),
        "wall_s": round(st["wall_s"], 2),
        "peak_state_gb": round(st["peak_bytes"] / 1e9, 2),
        "max_abs_mean_mass_drift": drift,
    }


def duct_batch(k=8, storage=1, steps=200, windows=3, seed=5):
    """The served workload's kernels without the loop: the first k duct
    requests of the seeded stream (FP16 storage as served, 8x8x8 tiles) in
    one batch, stepped with hipGraph replays; median of `windows` HIP-event
    windows of `steps` steps. The serving capacity of this mix."""
    import mrlbm
    from mrlbm import serve

    reqs = list(serve.duct_requests(k, seed=seed))
    lats = []
    for r in reqs:
        lat = mrlbm.Lattice((r.nx, r.ny, r.nz), q=19, tau=0.5 + 3.0 * r.nu, walls=(0, 1, 1), force=(r.gx, 0, 0),
                            tile=(8, 8, 8), storage=storage, seed=r.seed)
        sp = [r.sphere[i] for i in range(r.nspheres)]
        lat.set_geometry(centres=[(s[0], s[1], s[2]) for s in sp], radii=[s[3] for s in sp])
        lat.init_uniform()
        lats.append(lat)
    nodes = sum(r.nx * r.ny * r.nz for r in reqs)
    b = mrlbm.Batch(q=19, tile=(8, 8, 8), storage=storage)
    for lat in lats:
        b.add(lat)
    b.step(50, graph_steps=50)
    per = sorted(b.time_steps(steps, graph_steps=50) / steps for _ in range(windows))
    b.close()
    for lat in lats:
        lat.close()
    return {
        "workload": f"first {k} requests of the served duct stream (seed {seed}) in one batch, "
                    f"{['FP32', 'FP16', 'BF16'][storage]} moments, 8x8x8 tiles, graphs of 50 steps",
        "nodes": nodes,
        "value": round(nodes / (per[len(per) // 2] * 1e-3) / 1e6, 1),
        "unit": "MLUPS",
        "step_ms_windows": [round(t, 4) for t in per],
    }


def launch_windows(lat, launches, windows=3):
    """Mean launch time of `lat`'s step in seconds: the median over `windows`
    windows of `launches` back-to-back launches, each timed with HIP events
    on the lattice's stream."""
    per = sorted(lat.time_steps(launches) * 1e-3 / launches for _ in range(windows))
    return per[len(per) // 2], per


def reduced_precision_variants(n, tau, u0, launches, warmup):
    """Same workload with FP32, FP16 and BF16 moment storage (FP32 arithmetic
    in all three), each timed the same way: `warmup` launches, then the
    median of three event-timed windows of `launches` launches. The speedups
    divide matched windows. Each reduced-precision run's fields are compared
    with the FP32-storage run after the same number of steps (per-field
    relative L2 of rho - 1, rho and u)."""
    import numpy as np
    import mrlbm

    res, ref = {}, None
    for name, st in (("fp32", mrlbm.STORE_FP32), ("fp16", mrlbm.STORE_FP16), ("bf16", mrlbm.STORE_BF16)):
        lat = mrlbm.Lattice((n, n, n), q=19, tau=tau, storage=st)
        lat.init_taylor_green(u0)
        info = lat.info()
        lat.step(warmup)
        lat.synchronize()
        avg, per = launch_windows(lat, launches)
        rho, ux, uy, uz = lat.macroscopic()
        lat.close()
        achieved = info["bytes_per_lu"] * n ** 3 / avg / 1e9
        res[name] = {
            "value": round(n ** 3 / avg / 1e6, 1),
            "unit": "MLUPS",
            "timing": f"median of 3 windows of {launches} launches after {warmup} warm-up launches (HIP events)",
            "window_launch_us": [round(t * 1e6, 2) for t in per],
            "tile": [info["bx"], info["by"], info["bz"]],
            "bytes_per_lu": info["bytes_per_lu"],
            "mean_launch_us": round(avg * 1e6, 2),
            "achieved_gbs": round(achieved, 1),
            "roofline_frac": round(achieved / PEAK_HBM_GBS, 4),
        }
        if ref is None:
            ref = (avg, rho, np.stack([ux, uy, uz]))
            continue

        def rl2(a, b):
            return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

        res[name]["speedup_vs_fp32"] = round(ref[0] / avg, 3)
        res[name]["rel_l2_vs_fp32_storage"] = {"rho-1": rl2(rho - 1.0, ref[1] - 1.0), "rho": rl2(rho, ref[1]),
                                               "u": rl2(np.stack([ux, uy, uz]), ref[2])}
    return res


def channel_d3q27(steps, shape=(512, 256, 256), g=1e-6, nu=0.05, storages=("fp32", "fp16", "bf16")):
    """D3Q27 force-driven duct (x periodic, walls on y and z) from rest at each
    storage precision: throughput and the x-averaged streamed velocity against
    the transient rectangular-duct series at t = steps (and against the
    first storage's run, FP32 by default)."""
    import numpy as np
    import mrlbm
    from mrlbm import analytic, geometry

    types = geometry.classify(shape, (0, 1, 1))
    ser = analytic.duct_start_up(shape[1], shape[2], g, nu, t=steps)
    out, ref = {}, None
    codes = {"fp32": mrlbm.STORE_FP32, "fp16": mrlbm.STORE_FP16, "bf16": mrlbm.STORE_BF16}
    for name in storages:
        st = codes[name]
        with mrlbm.Lattice(shape, q=27, tau=0.5 + 3 * nu, walls=(0, 1, 1), force=(g, 0, 0), storage=st) as lat:
            lat.set_node_types(types)
            lat.init_uniform()
            lat.synchronize()
            tail = min(1000, steps // 2)
            t0 = time.perf_counter()
            ms = lat.time_steps(steps - tail)
            lat.velocity_residual()  # reference velocity `tail` steps before the end
            ms += lat.time_steps(tail)
            lat.synchronize()
            wall = time.perf_counter() - t0
            residual = lat.velocity_residual()
            info = lat.info()
            ux = lat.moments()[1].reshape(shape, order="F").mean(axis=0) / 3.0 - g  # stored u carries +F/rho
        if ref is None:
            ref = ux
        n = float(np.prod(shape))
        out[name] = {
            "value": round(n * steps / wall / 1e6, 1),
            "unit": "MLUPS",
            "mean_launch_us": round(ms * 1e3 / steps, 2),
            "bytes_per_lu": info["bytes_per_lu"],
            "velocity_residual_last_steps": [tail, residual],
            "rel_l2_vs_duct_series": float(np.linalg.norm(ux - ser) / np.linalg.norm(ser)),
            "rel_l2_vs_fp32_storage": float(np.linalg.norm(ux - ref) / np.linalg.norm(ref)),
        }
    return {"workload": f"D3Q27 duct {shape[0]}x{shape[1]}x{shape[2]}, x periodic, y/z walls, gx={g}, "
                        f"nu={nu}, from rest, {steps} steps", **out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--tile", default="", help="bx,by,bz (default: the library's choice)")
    ap.add_argument("--kernel", default="auto", choices=["auto", "tile", "column"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true", help="skip the FP16/BF16 storage lines")
    ap.add_argument("--storage", default="fp32", choices=["fp32", "fp16", "bf16"],
                    help="moment storage of the headline lattice (arithmetic is FP32 in all)")
    ap.add_argument("--batch-steps", type=int, default=2000, help="steps of each lattice in the batching comparison")
    ap.add_argument("--channel-steps", type=int, default=20000, help="steps of the D3Q27 duct at each precision")
    ap.add_argument("--no-channel", action="store_true", help="skip the D3Q27 duct lines")
    ap.add_argument("--channel-storage", default="fp32,fp16,bf16",
                    help="comma-separated storage formats of the D3Q27 duct lines (FP32 first is the reference)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-cavity batching comparison")
    ap.add_argument("--no-serve", action="store_true", help="skip the served duct stream")
    ap.add_argument("--duct-batch", type=int, default=8,
                    help="batch the first K served duct requests and time the kernels alone (0: skip)")
    ap.add_argument("--duct-batch-only", action="store_true",
                    help="print only the duct-batch timing (for profiling the served mix)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the lattice step is single-device in this build", file=sys.stderr)
        sys.exit(2)

    import mrlbm

    if args.duct_batch_only:
        print(json.dumps({"duct_batch": duct_batch(max(1, args.duct_batch))}), flush=True)
        return

    n = args.n
    tile = tuple(int(v) for v in args.tile.split(",")) if args.tile else None
    kernel = {"auto": mrlbm.KERNEL_AUTO, "tile": mrlbm.KERNEL_TILE, "column": mrlbm.KERNEL_COLUMN}[args.kernel]
    nu, u0 = 0.01, 0.05
    tau = 0.5 + 3 * nu
    N = n ** 3
    print(f"bench: D3Q19 TGV {n}^3 {args.storage} kernel {args.kernel} tile {tile} steps {args.steps} "
          f"warmup {args.warmup}",
          file=sys.stderr, flush=True)

    copy_gbs, _ = mrlbm.copy_bandwidth(1 << 30, 20)

    storage = {"fp32": mrlbm.STORE_FP32, "fp16": mrlbm.STORE_FP16, "bf16": mrlbm.STORE_BF16}[args.storage]
    lat = mrlbm.Lattice((n, n, n), q=19, tau=tau, tile=tile, kernel=kernel, storage=storage)
    lat.init_taylor_green(u0)
    info = lat.info()
    tile = (info["bx"], info["by"], info["bz"])
    kname = {mrlbm.KERNEL_TILE: "tile", mrlbm.KERNEL_COLUMN: "column"}[info["kernel"]]
    # roofline of the step kernel: median of three event-timed windows of at
    # least 200 launches (untimed by the headline), then the headline's own
    # W warm-up steps and exactly K timed steps
    roof_launches = max(200, args.steps)
    avg_launch_s, roof_windows = launch_windows(lat, roof_launches)
    lat.step(args.warmup)
    lat.synchronize()
    t0 = time.perf_counter()
    ev_ms = lat.time_steps(args.steps)  # HIP events on the lattice stream around the K launches
    lat.synchronize()
    t1 = time.perf_counter()
    mass, ke = lat.reduce()
    if not (ke == ke) or ke <= 0:
        raise RuntimeError("lattice state is not finite after the timed run")
    lat.close()

    wall = t1 - t0
    mlups = N * args.steps / wall / 1e6
    b_lu = info["bytes_per_lu"]
    achieved = b_lu * N / avg_launch_s / 1e9
    key = f"tgv{n}_d3q19_{args.storage}_{kname}{tile[0]}x{tile[1]}x{tile[2]}"
    traffic = traffic_from_profiles(key)

    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(mlups / REF_PUBLISHED_MLUPS, 3),
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"D3Q19 Taylor-Green vortex {n}^3 triply periodic, {args.storage.upper()} moment storage, "
                        f"u0={u0}, nu={nu} (tau={tau}), {kname} kernel {tile[0]}x{tile[1]}x{tile[2]}",
            "lattice": [n, n, n],
            "moment_storage": args.storage,
            "tile": list(tile),
            "kernel": kname,
            "bytes_per_lu": b_lu,
            "vs_baseline_ref": "A100 11390 MLUPS, D3Q19 FP32 256^3 (reference README.md:141)",
            "copy_kernel_gbs": round(copy_gbs, 1),
            "mean_launch_us": round(avg_launch_s * 1e6, 2),
            "roofline_timing": f"median of 3 HIP-event windows of {roof_launches} launches, before the timed steps",
            "roofline_window_launch_us": [round(t * 1e6, 2) for t in roof_windows],
            "timed_steps_event_launch_us": round(ev_ms * 1e3 / args.steps, 2),
            "final_mean_ke": ke,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic,
        },
    }
    def phase(name):
        print(f"bench: {name}", file=sys.stderr, flush=True)

    phase(f"headline {mlups:.1f} MLUPS")
    if not args.no_batch:
        phase("64-cavity batch")
        out["batching"] = batching_cavities(steps=args.batch_steps)
    if not args.no_serve:
        phase("served duct stream")
        out["serving"] = serving_ducts()
    if args.duct_batch > 0:
        phase("served duct mix, kernels alone")
        out["duct_batch"] = duct_batch(args.duct_batch)
    if not args.no_variants:
        phase("FP16/BF16 storage")
        out["variants"] = reduced_precision_variants(n, tau, u0, max(200, args.steps), max(20, args.warmup))
    if not args.no_channel:
        phase("D3Q27 duct at FP32/FP16/BF16")
        out["channel_d3q27"] = channel_d3q27(args.channel_steps, storages=args.channel_storage.split(","))
    if not args.no_cpu_baseline:
        phase("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
