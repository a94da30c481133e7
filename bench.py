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


def cpu_baseline(target_s):
    """FP64 oracle (OpenMP over the host cores) on a TGV 64^3 sample of the
    same workload, sized to about target_s seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle

    n = 64
    threads = _oracle.lib().orc_get_threads()
    o = _oracle.Oracle(n, n, n, q=19, tau=0.53)
    o.init_taylor_green(0.05)
    t1 = o.step_timed(2)
    steps = max(2, min(2000, int(target_s / max(t1 / 2, 1e-6))))
    t = o.step_timed(steps)
    return {
        "value": n ** 3 * steps / t / 1e6,
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"FP64 oracle, D3Q19 TGV {n}^3, {steps} steps ({t:.1f} s), OpenMP {threads} threads",
    }


def reduced_precision_variants(n, tau, u0, steps, warmup):
    """Same workload with FP16 and BF16 moment storage (FP32 arithmetic)."""
    import mrlbm

    res = {}
    for name, st in (("fp16", mrlbm.STORE_FP16), ("bf16", mrlbm.STORE_BF16)):
        lat = mrlbm.Lattice((n, n, n), q=19, tau=tau, storage=st)
        lat.init_taylor_green(u0)
        info = lat.info()
        lat.step(warmup)
        lat.synchronize()
        t0 = time.perf_counter()
        ev_ms = lat.time_steps(steps)
        lat.synchronize()
        t1 = time.perf_counter()
        lat.close()
        avg = ev_ms * 1e-3 / steps
        res[name] = {
            "value": round(n ** 3 * steps / (t1 - t0) / 1e6, 1),
            "unit": "MLUPS",
            "steps": steps,
            "tile": [info["bx"], info["by"], info["bz"]],
            "bytes_per_lu": info["bytes_per_lu"],
            "mean_launch_us": round(avg * 1e6, 2),
            "achieved_gbs": round(info["bytes_per_lu"] * n ** 3 / avg / 1e9, 1),
        }
    return res


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
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the lattice step is single-device in this build", file=sys.stderr)
        sys.exit(2)

    import mrlbm

    n = args.n
    tile = tuple(int(v) for v in args.tile.split(",")) if args.tile else None
    kernel = {"auto": mrlbm.KERNEL_AUTO, "tile": mrlbm.KERNEL_TILE, "column": mrlbm.KERNEL_COLUMN}[args.kernel]
    nu, u0 = 0.01, 0.05
    tau = 0.5 + 3 * nu
    N = n ** 3
    print(f"bench: D3Q19 TGV {n}^3 kernel {args.kernel} tile {tile} steps {args.steps} warmup {args.warmup}",
          file=sys.stderr, flush=True)

    copy_gbs, _ = mrlbm.copy_bandwidth(1 << 30, 20)

    lat = mrlbm.Lattice((n, n, n), q=19, tau=tau, tile=tile, kernel=kernel)
    lat.init_taylor_green(u0)
    info = lat.info()
    tile = (info["bx"], info["by"], info["bz"])
    kname = {mrlbm.KERNEL_TILE: "tile", mrlbm.KERNEL_COLUMN: "column"}[info["kernel"]]
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
    avg_launch_s = ev_ms * 1e-3 / args.steps
    b_lu = info["bytes_per_lu"]
    achieved = b_lu * N / avg_launch_s / 1e9
    key = f"tgv{n}_d3q19_fp32_{kname}{tile[0]}x{tile[1]}x{tile[2]}"
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
            "workload": f"D3Q19 Taylor-Green vortex {n}^3 triply periodic, FP32 moment storage, "
                        f"u0={u0}, nu={nu} (tau={tau}), {kname} kernel {tile[0]}x{tile[1]}x{tile[2]}",
            "lattice": [n, n, n],
            "tile": list(tile),
            "kernel": kname,
            "bytes_per_lu": b_lu,
            "vs_baseline_ref": "A100 11390 MLUPS, D3Q19 FP32 256^3 (reference README.md:141)",
            "copy_kernel_gbs": round(copy_gbs, 1),
            "mean_launch_us": round(avg_launch_s * 1e6, 2),
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
    if not args.no_variants:
        out["variants"] = reduced_precision_variants(n, tau, u0, args.steps // 2, args.warmup)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
