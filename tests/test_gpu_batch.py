"""Batched stepping (one launch per step over many lattices) must give the
same bits as stepping every lattice alone."""
import numpy as np
import pytest

import mrlbm
from mrlbm import geometry

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


CASES = [  # (n, tau, lid velocity)
    (32, 0.56, 0.05), (48, 0.62, 0.04), (64, 0.55, 0.05), (32, 0.9, 0.03), (16, 0.7, 0.05),
]


def make(n, tau, u, storage=mrlbm.STORE_FP32, seed=0):
    lat = mrlbm.Lattice((n, n, n), q=19, tau=tau, walls=(1, 1, 1), tile=(8, 8, 8), kernel=mrlbm.KERNEL_TILE,
                        storage=storage, seed=seed)
    lat.set_bc_velocity(0, (u, 0, 0))
    lat.set_node_types(geometry.cavity(n))
    lat.init_uniform()
    return lat


def alone(steps, storage=mrlbm.STORE_FP32):
    out = []
    for i, (n, tau, u) in enumerate(CASES):
        with make(n, tau, u, storage, seed=i) as lat:
            lat.step(steps)
            out.append(lat.moments())
    return out


@pytest.mark.parametrize("mode", ["plain", "graph", "persistent"])
def test_batch_equals_alone(mode):
    steps = 40
    ref = alone(steps)
    with mrlbm.Batch(q=19, tile=(8, 8, 8)) as b:
        lats = [make(n, tau, u, seed=i) for i, (n, tau, u) in enumerate(CASES)]
        for lat in lats:
            b.add(lat)
        if mode == "persistent":
            b.set_grid(100)  # below both launches' item counts (296 bulk, 568 boundary): ranged kernel
        b.step(steps, graph_steps=10 if mode == "graph" else 0)
        for lat, r in zip(lats, ref):
            assert lat.info()["step"] == steps
            assert np.array_equal(lat.moments(), r)
        for lat in lats:
            lat.close()


@pytest.mark.parametrize("grid", [0, 100])
def test_batch_fp16_equals_alone(grid):
    steps = 30
    ref = alone(steps, mrlbm.STORE_FP16)
    with mrlbm.Batch(q=19, tile=(8, 8, 8), storage=mrlbm.STORE_FP16) as b:
        lats = [make(n, tau, u, mrlbm.STORE_FP16, seed=i) for i, (n, tau, u) in enumerate(CASES)]
        for lat in lats:
            b.add(lat)
        b.set_grid(grid)
        b.step(steps, graph_steps=10)
        for lat, r in zip(lats, ref):
            assert np.array_equal(lat.moments(), r)


def test_batch_membership_changes():
    """Lattices join and leave between steps (continuous batching) and a
    member is re-initialised mid-run; each still matches its lone run."""
    n0, tau0, u0 = CASES[0]
    n1, tau1, u1 = CASES[1]
    with mrlbm.Batch(q=19, tile=(8, 8, 8)) as b:
        a = make(n0, tau0, u0, seed=0)
        b.add(a)
        b.step(20, graph_steps=10)
        c = make(n1, tau1, u1, seed=1)
        b.add(c)  # joins at a's step 20
        b.step(20, graph_steps=10)
        b.remove(a)
        b.step(10)
        a_m, c_m = a.moments(), c.moments()
        a.init_uniform()  # re-initialise while detached, then rejoin
        b.add(a)
        b.step(10)
        a2 = a.moments()
    with make(n0, tau0, u0, seed=0) as r:
        r.step(40)
        assert np.array_equal(a_m, r.moments())
    with make(n1, tau1, u1, seed=1) as r:
        r.step(30)
        assert np.array_equal(c_m, r.moments())
    with make(n0, tau0, u0, seed=0) as r:
        r.step(10)
        assert np.array_equal(a2, r.moments())


def test_member_cannot_step_alone():
    with mrlbm.Batch() as b:
        lat = make(16, 0.7, 0.05)
        b.add(lat)
        with pytest.raises(mrlbm.MrlbmError):
            lat.step(1)
        b.remove(lat)
        lat.step(1)
        lat.close()


def test_autotune_picks_fitting_shape():
    dims = [(32, 32, 32), (48, 48, 48), (64, 64, 64)]
    best, ms = mrlbm.autotune(dims, steps=20)
    say("autotune", best, ms)
    assert ms[(32, 4, 4)] is None and ms[(64, 4, 2)] is None  # 48 is not a multiple of 32 or 64
    assert best in ((8, 8, 8), (16, 8, 4))


def test_autotune_with_walls_times_boundary_tiles():
    """Wall lattices are tuned with device-classified domain walls (the
    boundary body runs); the choice is a fitting shape and every fitting
    shape is timed."""
    dims = [(32, 32, 32), (48, 48, 48)]
    best, ms = mrlbm.autotune(dims, steps=20, walls=[(1, 1, 1), (0, 1, 1)])
    say("autotune walls", best, ms)
    assert ms[(32, 4, 4)] is None and ms[(64, 4, 2)] is None
    assert best in ((8, 8, 8), (16, 8, 4))
    assert all(ms[s] > 0 for s in ((8, 8, 8), (16, 8, 4)))
    with pytest.raises(ValueError):
        mrlbm.autotune(dims, steps=20, walls=[(1, 1, 1)])
