"""Serving with continuous batching: every result is the same alone and in
the mix; the loop reports latency percentiles and throughput."""
import numpy as np
import pytest

import mrlbm
from mrlbm import geometry, serve

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def small_stream(n, rate, seed):
    reqs = serve.duct_requests(n, rate=rate, seed=seed)
    for i, r in enumerate(reqs):  # shrink to test sizes, keep the rest of the draw
        r.nx, r.ny, r.nz = 64, 32, 32
        r.nspheres = min(r.nspheres, 3)
        for k in range(r.nspheres):
            r.sphere[k][0] %= 64
            r.sphere[k][1] %= 32
            r.sphere[k][2] %= 32
            r.sphere[k][3] = min(r.sphere[k][3], 5.0)
        r.steps = 100 + 37 * i
    return reqs


def alone(r, storage):
    lat = mrlbm.Lattice((r.nx, r.ny, r.nz), q=19, tau=0.5 + 3 * r.nu, walls=(0, 1, 1), force=(r.gx, 0, 0),
                        tile=(8, 8, 8), kernel=mrlbm.KERNEL_TILE, storage=storage, seed=r.seed)
    solid = geometry.spheres((r.nx, r.ny, r.nz), [tuple(r.sphere[k][:3]) for k in range(r.nspheres)],
                             [r.sphere[k][3] for k in range(r.nspheres)])
    lat.set_node_types(geometry.classify((r.nx, r.ny, r.nz), (0, 1, 1), solid))
    lat.init_uniform()
    lat.step(r.steps)
    m = lat.moments()
    lat.close()
    return m


@pytest.mark.parametrize("storage", [mrlbm.STORE_FP32, mrlbm.STORE_FP16])
def test_served_results_equal_lone_runs(storage):
    reqs = small_stream(6, 1000.0, 3)
    bufs = [serve.attach_output(r) for r in reqs]
    st = serve.run(reqs, storage=storage, quantum=50, graph_steps=25, realtime=False)
    say("stats", st)
    assert all(r.status == 0 for r in reqs)
    for r, b in zip(reqs, bufs):
        assert np.array_equal(b.reshape(10, -1), alone(r, storage))
        assert abs(r.mass1 - r.mass0) < 1e-3


def test_realtime_stream_stats():
    reqs = small_stream(16, 40.0, 9)
    st = serve.run(reqs, storage=mrlbm.STORE_FP16, quantum=100, graph_steps=50, realtime=True)
    say("stats", st)
    lat = sorted(r.latency_s for r in reqs)
    assert all(r.status == 0 for r in reqs)
    assert 0 < st["p50_s"] <= st["p99_s"] <= st["max_s"] + 1e-12
    assert abs(st["max_s"] - lat[-1]) < 1e-9
    assert st["mlups"] > 0 and st["lattice_updates"] == sum(r.nx * r.ny * r.nz * r.steps for r in reqs)
    # requests overlapped: fewer rounds than the serial sum of rounds
    assert st["rounds"] < sum(-(-r.steps // 100) for r in reqs)


def test_failed_requests_leave_the_others_unchanged():
    """One request larger than the budget (ENOMEM), one with extents off the
    tile (EINVAL) and one driven by a NaN force, whose lattice turns
    non-finite in its first round (ENUMERIC, found by the per-round FP64
    check and released): the others finish bit-identical to lone runs."""
    reqs = small_stream(6, 1000.0, 5)
    reqs[1].nx, reqs[1].ny, reqs[1].nz = 256, 128, 128  # 4.2 M nodes, ~214 MB: over the 0.1 GB budget
    reqs[2].nx = 70
    reqs[4].gx = float("nan")
    bufs = [serve.attach_output(r) if i not in (1, 2) else None for i, r in enumerate(reqs)]
    st = serve.run(reqs, storage=mrlbm.STORE_FP16, budget_gb=0.1, quantum=50, graph_steps=25, realtime=False)
    say("stats", st, [r.status for r in reqs])
    assert [r.status for r in reqs] == [0, -3, -1, 0, -6, 0]
    assert st["failed"] == 3
    for i in (0, 3, 5):
        assert np.array_equal(bufs[i].reshape(10, -1), alone(reqs[i], mrlbm.STORE_FP16)), i
        assert np.isfinite(reqs[i].mass1)
