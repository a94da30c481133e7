"""The serving loop on a simulated device (mrlbm_serve_simulate): the same
admission / round / retirement code as the GPU loop, with a virtual clock,
so its scheduling can be checked exactly on the CPU."""
import numpy as np
import pytest

from mrlbm import _lib, serve


def state_bytes(r, q=19, storage=0, tile=(8, 8, 8)):
    n = r.nx * r.ny * r.nz
    qf = 5 if q == 19 else 9
    ghosts = 2.0 * 2.0 * qf * 4.0 * n * sum(1.0 / t for t in tile)
    return n * (10.0 * (4 if storage == 0 else 2) + 1.0) + ghosts


def two_requests(steps_a, steps_b, arrival_b=0.0):
    reqs = serve.duct_requests(2, rate=10.0, seed=1)
    for r in reqs:
        r.nx, r.ny, r.nz = 64, 32, 32
        r.arrival_s = 0.0
    reqs[0].steps, reqs[1].steps = steps_a, steps_b
    reqs[1].arrival_s = arrival_b
    return reqs


def test_two_requests_exact_timeline():
    """A (150 steps) and B (100 steps) arrive together; rounds of at most 100
    steps never pass a member's budget: round 1 = 100 steps of both, B leaves;
    round 2 = 50 steps of A alone. On a device doing R updates per second,
    B finishes at 200 N / R and A at 250 N / R."""
    reqs = two_requests(150, 100)
    rate = 1000.0  # MLUPS
    st = serve.simulate(reqs, rate, quantum=100, graph_steps=50)
    n = 64 * 32 * 32
    R = rate * 1e6
    assert reqs[1].latency_s == pytest.approx(200 * n / R, rel=1e-12)
    assert reqs[0].latency_s == pytest.approx(250 * n / R, rel=1e-12)
    assert st["rounds"] == 2
    assert st["lattice_updates"] == n * (150 + 100)
    assert st["busy_s"] == pytest.approx(250 * n / R, rel=1e-12)


def test_late_arrival_joins_at_round_boundary():
    """B arrives during A's first round: it is admitted when that round ends
    and joins the next one; A's rounds are not cut short by the arrival."""
    n = 64 * 32 * 32
    R = 1000.0 * 1e6
    t_round = 100 * n / R
    reqs = two_requests(300, 100, arrival_b=0.5 * t_round)
    serve.simulate(reqs, 1000.0, quantum=100, graph_steps=50)
    assert reqs[1].start_s == pytest.approx(t_round, rel=1e-12)
    # B joins round 2 (both run 100 steps: 2 t_round), then A runs round 3 alone
    assert reqs[1].finish_s == pytest.approx(t_round + 2 * t_round, rel=1e-12)
    assert reqs[0].finish_s == pytest.approx(reqs[1].finish_s + t_round, rel=1e-12)


@pytest.mark.parametrize("budget_gb,max_batch", [(200.0, 0), (1.0, 0), (200.0, 4)])
def test_stream_invariants(budget_gb, max_batch):
    """Over a seeded stream: every request completes after exactly its step
    budget (total updates = sum of N * steps), never starts before it
    arrives, the admitted state never exceeds the budget, and latency is
    finish - arrival."""
    reqs = serve.duct_requests(300, rate=20.0, seed=7)
    st = serve.simulate(reqs, 40000.0, admit_s=0.002, budget_gb=budget_gb, quantum=100, max_batch=max_batch)
    total = sum(r.nx * r.ny * r.nz * r.steps for r in reqs)
    assert st["lattice_updates"] == total
    assert st["peak_bytes"] <= budget_gb * 1e9
    for r in reqs:
        assert r.status == 0
        assert r.start_s >= r.arrival_s - 1e-12
        assert r.finish_s >= r.start_s
        assert r.latency_s == pytest.approx(r.finish_s - r.arrival_s, abs=1e-12)
    lat = sorted(r.latency_s for r in reqs)
    assert st["p50_s"] == pytest.approx(np.percentile(lat, 50), rel=1e-12)
    assert st["p99_s"] == pytest.approx(np.percentile(lat, 99), rel=1e-12)
    assert st["max_s"] == pytest.approx(lat[-1])
    assert st["busy_s"] <= st["wall_s"] + 1e-9


def test_limits_keep_work_and_makespan():
    """A budget that admits few lattices at a time, or a batch limit, changes
    the order of service but not the work: on a device whose rate does not
    depend on the batch, the total updates, busy time and makespan are the
    same. (Under overload the limits even lower the mean latency, service
    then being closer to first-come-first-served than to equal sharing.)"""
    out = []
    for kw in (dict(budget_gb=200.0), dict(budget_gb=0.5), dict(max_batch=2)):
        reqs = serve.duct_requests(200, rate=20.0, seed=9)
        st = serve.simulate(reqs, 40000.0, **kw)
        out.append(st)
        if "budget_gb" in kw:
            assert st["peak_bytes"] <= kw["budget_gb"] * 1e9
    for st in out[1:]:
        assert st["lattice_updates"] == out[0]["lattice_updates"]
        assert st["busy_s"] == pytest.approx(out[0]["busy_s"], rel=1e-9)
        assert st["wall_s"] == pytest.approx(out[0]["wall_s"], rel=1e-6)


def test_batch_limit_respected():
    """With max_batch = 3, no more than 3 requests are ever between start and
    finish at once."""
    reqs = serve.duct_requests(120, rate=50.0, seed=11)
    serve.simulate(reqs, 40000.0, max_batch=3)
    ev = sorted([(r.start_s, 1) for r in reqs] + [(r.finish_s, -1) for r in reqs], key=lambda e: (e[0], e[1]))
    live = peak = 0
    for _, d in ev:
        live += d
        peak = max(peak, live)
    assert peak <= 3


def test_request_over_budget_fails_alone():
    """A request larger than the whole budget fails (MRLBM_ENOMEM) at its
    arrival; the one that fits is served as if it were alone."""
    reqs = two_requests(10, 10)
    reqs[0].nx = 256  # 256 x 32 x 32 does not fit 0.8 MB; 64 x 32 x 32 does
    st = serve.simulate(reqs, 1000.0, budget_gb=state_bytes(reqs[1]) * 1.01e-9)
    assert reqs[0].status == -3 and reqs[1].status == 0
    assert st["failed"] == 1
    assert st["lattice_updates"] == 64 * 32 * 32 * 10
    assert st["p50_s"] == pytest.approx(reqs[1].latency_s)


def test_failures_are_per_request():
    """Among 40 requests: one whose extents are not a multiple of the tile
    (EINVAL), one with nine spheres (EINVAL), one with nu = 0 (EINVAL) and
    one driven by a non-finite force, which turns non-finite in its first
    round (ENUMERIC) and leaves the batch then. Every other request runs
    exactly its step budget; the work done is theirs plus one round of the
    non-finite one; percentiles count completed requests only."""
    reqs = serve.duct_requests(40, rate=20.0, seed=13)
    reqs[3].nx = 70
    reqs[7].nspheres = 9
    reqs[11].nu = 0.0
    reqs[17].gx = float("nan")
    st = serve.simulate(reqs, 40000.0, admit_s=0.001, quantum=100, max_batch=6)
    bad = {3: -1, 7: -1, 11: -1, 17: -6}
    for i, r in enumerate(reqs):
        assert r.status == bad.get(i, 0), i
    assert st["failed"] == 4
    ok = [r for i, r in enumerate(reqs) if i not in bad]
    first_round = min(100, reqs[17].steps)
    total = sum(r.nx * r.ny * r.nz * r.steps for r in ok) + reqs[17].nx * reqs[17].ny * reqs[17].nz * first_round
    assert st["lattice_updates"] == total
    lat = sorted(r.latency_s for r in ok)
    assert st["p50_s"] == pytest.approx(np.percentile(lat, 50), rel=1e-12)
    assert st["max_s"] == pytest.approx(lat[-1])
    # the non-finite request left after one round: its finish follows its start by one round
    assert reqs[17].finish_s > reqs[17].start_s


def test_unordered_stream_is_rejected():
    reqs = two_requests(10, 10)
    reqs[0].arrival_s, reqs[1].arrival_s = 1.0, 0.5
    with pytest.raises(_lib.MrlbmError):
        serve.simulate(reqs, 1000.0)
